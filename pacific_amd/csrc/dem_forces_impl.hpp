// Force sweep launchers of one precision and body type: the translation unit
// defines DEM_FORCES_MIXED (0 or 1) and DEM_FORCES_SHAPES (0 or 1) and
// includes this file, so that the heavy sweep instantiations compile in
// parallel.
#include "dem_forces.hpp"

namespace dem {

#if DEM_FORCES_SHAPES
template <int MODE, bool MIXED>
void launch_forces_shp(const Particles& P, const Sorted& S, const Grid& G, const BoxObstacle* obs,
                       const HookeModel* models, const StepParams& sp, DevStatus* st, int* pairs, const NbrList& L,
                       const VerletState& V, const NbrList& CL, hipStream_t stream) {
  // the contacts come from the narrow phase's records: one-wave workgroups
  hipLaunchKernelGGL((k_forces<MODE, MIXED, true, true>), dim3(nblk_f(sp.n, kSphereBlock)), dim3(kSphereBlock), 0,
                     stream, P, S, G, obs, models, sp, st, pairs, L, V, CL);
}
template void launch_forces_shp<kForcesOnly, DEM_FORCES_MIXED>(const Particles&, const Sorted&, const Grid&,
    const BoxObstacle*, const HookeModel*, const StepParams&, DevStatus*, int*, const NbrList&, const VerletState&,
    const NbrList&, hipStream_t);
template void launch_forces_shp<kForcesKick, DEM_FORCES_MIXED>(const Particles&, const Sorted&, const Grid&,
    const BoxObstacle*, const HookeModel*, const StepParams&, DevStatus*, int*, const NbrList&, const VerletState&,
    const NbrList&, hipStream_t);
template void launch_forces_shp<kFused, DEM_FORCES_MIXED>(const Particles&, const Sorted&, const Grid&,
    const BoxObstacle*, const HookeModel*, const StepParams&, DevStatus*, int*, const NbrList&, const VerletState&,
    const NbrList&, hipStream_t);
#else
template <int MODE, bool MIXED>
void launch_forces_sph(const Particles& P, const Sorted& S, const Grid& G, const BoxObstacle* obs,
                       const HookeModel* models, const StepParams& sp, DevStatus* st, int* pairs, const NbrList& L,
                       const VerletState& V, const NbrList& CL, hipStream_t stream) {
  // one wave per workgroup: at 1-3 waves per SIMD, whole-wave workgroups
  // fill the CUs more evenly than 4-wave ones (the mixed sphere sweep of the
  // 1M gas: 92.7 us with 256 threads, 85.7 with 64); a launch that collects
  // statistics keeps 256 threads, the granularity of its partial sums
  const int bs = sp.stats ? 256 : kSphereBlock;
  const dim3 grid(nblk_f(sp.n, bs)), block(bs);
  if (sp.stats || sp.mem.cap || sp.fsum)
    hipLaunchKernelGGL((k_forces<MODE, MIXED, false, true>), grid, block, 0, stream, P, S, G, obs, models, sp, st,
                       pairs, L, V, CL);
  else if (!MIXED)
    hipLaunchKernelGGL((k_forces64<MODE>), grid, block, 0, stream, P, S, G, obs, models, sp, st, pairs, L, V, CL);
  else
    hipLaunchKernelGGL((k_forces<MODE, MIXED, false, false>), grid, block, 0, stream, P, S, G, obs, models, sp, st,
                       pairs, L, V, CL);
}
template void launch_forces_sph<kForcesOnly, DEM_FORCES_MIXED>(const Particles&, const Sorted&, const Grid&,
    const BoxObstacle*, const HookeModel*, const StepParams&, DevStatus*, int*, const NbrList&, const VerletState&,
    const NbrList&, hipStream_t);
template void launch_forces_sph<kForcesKick, DEM_FORCES_MIXED>(const Particles&, const Sorted&, const Grid&,
    const BoxObstacle*, const HookeModel*, const StepParams&, DevStatus*, int*, const NbrList&, const VerletState&,
    const NbrList&, hipStream_t);
template void launch_forces_sph<kFused, DEM_FORCES_MIXED>(const Particles&, const Sorted&, const Grid&,
    const BoxObstacle*, const HookeModel*, const StepParams&, DevStatus*, int*, const NbrList&, const VerletState&,
    const NbrList&, hipStream_t);
#endif

}  // namespace dem
