// Host entry points of the DEM kernels (dem_kernels.hip)
#pragma once
#include <hip/hip_runtime.h>

#include "dem_types.hpp"

namespace dem {

enum ForceMode { kModeForcesOnly = 0, kModeForcesKick = 1, kModeFused = 2 };

// first half kick + drift (or a full Move of the explicit schemes) in place,
// then cell key and arrival slot
hipError_t launch_move(const Particles& P, const Sorted& S, const Grid& G, const StepParams& sp, DevStatus* st,
                       double dtv, double dtx, const VerletState& V, hipStream_t stream);
// cell key and arrival slot of the current positions (no motion)
hipError_t launch_key(const Particles& P, const Sorted& S, const Grid& G, const StepParams& sp, DevStatus* st,
                      const VerletState& V, hipStream_t stream);
// counts -> start, scatter, per-cell ordering; cx = position buffer to gather;
// with gate, only when *gate == stamp
hipError_t launch_bin(const Particles& P, const Sorted& S, const Grid& G, int n, int cx, hipStream_t stream,
                      const int* gate = nullptr, int stamp = 0);
// Verlet list build (keys on its grid, binning, candidates); stamp > 0: only
// if the list was found stale (DevStatus::rebuild == stamp)
hipError_t launch_verlet_build(const Particles& P, const Sorted& S, const VerletState& V, int n, int cx,
                               DevStatus* st, const NbrList& L, int stamp, hipStream_t stream);
int scan_blocks(int ncells);
// zero n ints with a kernel (safe inside captured graphs, unlike hipMemsetAsync)
hipError_t launch_clear(int* p, int n, hipStream_t stream);
// contacts of this step from the Verlet list L into CL
hipError_t launch_vfilter(const Particles& P, const Sorted& S, const Grid& G, const VerletState& V, int n, int cx,
                          const NbrList& L, const NbrList& CL, DevStatus* st, hipStream_t stream);
// folds the per-block contact statistics of the last force launch over n
// particles into the running totals (6 doubles), at the given time
hipError_t launch_stats_fold(const double* part, int n, double* tot, double time, hipStream_t stream);
// neighbour candidates of every sorted particle into L
hipError_t launch_detect(const Sorted& S, const Grid& G, int n, const NbrList& L, const DevStatus* st,
                         hipStream_t stream, int* mark = nullptr);
// narrow phase of a system with non-spherical bodies: the contacts of every
// sorted particle (candidates of L, or its 27 cells when L.count is null)
// into sp.cf, and the contact statistics when sp.stats is set
hipError_t launch_narrow(const Particles& P, const Sorted& S, const Grid& G, const BoxObstacle* obs,
                         const StepParams& sp, DevStatus* st, const NbrList& L, hipStream_t stream);
// force sweep over the candidates of L (L.count == nullptr: 27-cell sweep)
hipError_t launch_forces(int mode, bool mixed, const Particles& P, const Sorted& S, const Grid& G,
                         const BoxObstacle* obs, const HookeModel* models, const StepParams& sp, DevStatus* st,
                         int* pairs, const NbrList& L, const VerletState& V, const NbrList& CL,
                         hipStream_t stream);

}  // namespace dem
