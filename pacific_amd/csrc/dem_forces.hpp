// Force sweep kernels of the DEM time step (shared by the per-precision,
// per-body-type translation units dem_forces_*.hip).
//
//   k_forces<MODE>       one thread per particle (storage order; sorted order
//                        with non-spherical bodies): the listed (or, with no
//                        list, the 27-cell) candidates with minimum-image
//                        shifts, closest points, Hooke-dashpot-Coulomb force
//                        and torque gathered in registers (no atomics), then
//                        the second half kick, optionally fused with the next
//                        step's first half kick and drift.
#pragma once
#include <hip/hip_runtime.h>

#include "dem_device.hpp"

namespace dem {

enum { kForcesOnly = 0, kForcesKick = 1, kFused = 2 };

// Second half kick of a sphere from its torsor (F, T), and inside a fused
// step the next Move, periodic swap and cell key (or Verlet staleness check).
template <int MODE>
__device__ __forceinline__ void sphere_tail(const Particles& P, const Sorted& S, const Grid& G, const StepParams& sp,
                                            DevStatus* st, const VerletState& VS, int s, int cx, int cv, Kin& kin,
                                            double mass, double crust, int id, double cf, double I, const d3& F,
                                            const d3& T) {
  const int nv = cv ^ 1;
  if (MODE == kForcesKick) {
    lf_kick(kin, F, T, mass, cf, I, 0.5 * sp.dt);
    st3(P.v[nv], s, kin.v);
    st3(P.w[nv], s, kin.w);
    st3(P.wb, s, kin.wb);
    return;
  }
  // kFused
  double depl;
  if (sp.scheme == kLeapFrog) {
    lf_kick(kin, F, T, mass, cf, I, 0.5 * sp.dt);
    depl = lf_move(kin, F, T, mass, cf, I, 0.5 * sp.dt, sp.dt);
  } else {
    AB2State ab;
    if (sp.scheme == kAdamsBashforth2) {
      ab.v_nm2 = ld3(P.v_nm2, s);
      ab.dUdt_nm2 = ld3(P.dUdt_nm2, s);
      ab.wb_nm2 = ld3(P.wb_nm2, s);
      ab.dWdt_nm2 = ld3(P.dWdt_nm2, s);
    }
    depl = other_move(sp.scheme, sp.first_step != 0, kin, ab, F, T, mass, cf, I, sp.dt, sp.dt);
    if (sp.scheme == kAdamsBashforth2) {
      st3(P.v_nm2, s, ab.v_nm2);
      st3(P.dUdt_nm2, s, ab.dUdt_nm2);
      st3(P.wb_nm2, s, ab.wb_nm2);
      st3(P.dWdt_nm2, s, ab.dWdt_nm2);
    }
  }
  if (depl > crust) raise_error(st, kMotionError, id, 0);
  wrap_and_bin(kin, G, S, st, s, id, 0, &VS);
  st3(P.x[cx ^ 1], s, kin.x);
  st3(P.v[nv], s, kin.v);
  st3(P.w[nv], s, kin.w);
  st3(P.wb, s, kin.wb);
  P.q[0][s] = kin.q.v.x;
  P.q[1][s] = kin.q.v.y;
  P.q[2][s] = kin.q.v.z;
  P.q[3][s] = kin.q.w;
}

// Force and torque of one body from the contacts the narrow phase recorded
// for its sorted slot k (ContactFeat), in the order it met them -- the order
// of the 27-cell sweep -- so that the sum is the one of an inline sweep.
template <bool MIXED, bool EXTRAS>
__device__ __forceinline__ void listed_contacts(const Particles& P, const Sorted& S, const BoxObstacle* __restrict__ obs,
                                                const HookeModel* __restrict__ models, const StepParams& sp,
                                                DevStatus* st, int* pairs, int cv, int k, int s, const Self& me, d3& F,
                                                d3& T) {
  using FV = ForceVisitor<MIXED, EXTRAS>;
  FV vis{P, S, obs, models, sp, st, pairs, cv, me, F, T, f3{0.f, 0.f, 0.f}, f3{0.f, 0.f, 0.f}};
  vis.s = s;
  const ContactFeat& cf = sp.cf;
  const int c = min(cf.count[k], cf.cap);
  // the slots are read 16 at a time with independent loads, and only those
  // holding a contact (not kNoFeat) are evaluated, in slot order
  for (int j0 = 0; j0 < c; j0 += 16) {
    unsigned live = 0u;
#pragma unroll
    for (int u = 0; u < 16; ++u)
      if (j0 + u < c && cf.ent[(size_t)(j0 + u) * cf.n + k] != kNoFeat) live |= 1u << u;
    while (live) {
      const int j = j0 + __builtin_ctz(live);
      live &= live - 1u;
      const size_t e = (size_t)j * cf.n + k;
      const int o = cf.ent[e];
      const d3 Pc{cf.f[3][e], cf.f[4][e], cf.f[5][e]}, ov{cf.f[6][e], cf.f[7][e], cf.f[8][e]};
      if (o >= 0) {
        if (o < sp.n) vis.pair(-1, o, cf.pid[e], d3{cf.f[0][e], cf.f[1][e], cf.f[2][e]}, 0., Pc, ov, -1.);
      } else if (-1 - o < sp.nobs) {
        vis.wall(-1 - o, Pc, ov, -1.);
      }
    }
  }
  F = vis.F;
  T = vis.T;
  if (MIXED) {
    F += d3{(double)vis.Fc.x, (double)vis.Fc.y, (double)vis.Fc.z};
    T += d3{(double)vis.Tc.x, (double)vis.Tc.y, (double)vis.Tc.z};
  }
  if (EXTRAS && sp.fsum) sp.fsum[s] = vis.fs;
  if (EXTRAS && sp.mem.cap) {
    sp.mem.n[sp.mem.cur ^ 1][s] = min(vis.nmem, sp.mem.cap);
    if (vis.nmem > sp.mem.cap) raise_error(st, kMemoryOverflow, me.id, vis.nmem);
  }
}

template <int MODE, bool MIXED, bool SHAPES, bool EXTRAS>
__device__ __forceinline__ void force_one(const Particles& P, const Sorted& S, const Grid& G,
                                          const BoxObstacle* __restrict__ obs, const HookeModel* __restrict__ models,
                                          const StepParams& sp, DevStatus* st, int* pairs, const NbrList& L,
                                          const VerletState& VS, const NbrList& CL, int t, StatAcc* acc) {
  const int cx = sp.cur & 1, cv = (sp.cur >> 1) & 1;
  const int s = SHAPES ? S.order[t] : t;
  const int k = SHAPES ? t : S.inv[t];
  // only after an interrupted binning
  if ((unsigned)s >= (unsigned)sp.n || (unsigned)k >= (unsigned)sp.n) return;
  Self me;
  if (VS.stamp) {
    // Verlet list: the sorted arrays hold the positions of the last build
    me.x = ld3(P.x[cx], s);
    me.r = P.rad[s];
  } else {
    const double* pr = S.posr + 4 * (size_t)k;
    me.x = d3{pr[0], pr[1], pr[2]};
    me.r = pr[3];
  }
  me.id = P.id[s];
  me.v = ld3(P.v[cv], s);
  me.w = ld3(P.w[cv], s);
  me.m = P.mass[s];
  me.c = P.crust[s];
  me.req = P.req[s];
  me.mat = P.mat[s];
  me.acc = acc;
  const double wf = P.wfac[s];
  // InitializeForce: torsor = weight at the centre (Particle.cpp:1408-1415)
  d3 F{sp.g[0] * wf, sp.g[1] * wf, sp.g[2] * wf};
  d3 T{0., 0., 0.};

  // non-spherical bodies: the basis (drift) and whether the body rotates
  // with the gyroscopic term; their contacts come from the narrow phase
  double B[SHAPES ? 9 : 1];
  bool gyro = false;
  if (SHAPES) {
    for (int j = 0; j < 9; ++j) B[j] = P.B[cx][j][s];
    const int sh = P.shape[s];
    gyro = sh >= 0 && P.shapes[sh].type != kShapeSphere;
    listed_contacts<MIXED, EXTRAS>(P, S, obs, models, sp, st, pairs, cv, k, s, me, F, T);
  } else {
    sweep_particle<MIXED, false, EXTRAS>(P, S, G, obs, sp.nobs, models, sp, st, pairs, 0, 0, sp.n, cv, k, me, F, T,
                                         nullptr, cx, &L, &VS, s, &CL);
  }
  // forces of the fluid (AppPRSHydroFT::ComputeForces, AppPRSHydroFT.cpp:
  // 27-44): an application after the contacts, the force at the centre
  if (sp.hyd) {
    const size_t n = (size_t)sp.n;
    F += d3{sp.hyd[s], sp.hyd[n + s], sp.hyd[2 * n + s]};
    T += d3{sp.hyd[3 * n + s], sp.hyd[4 * n + s], sp.hyd[5 * n + s]};
  }
  // inside a fused LeapFrog step the torsor is consumed here; it is stored by
  // the force-only and kick launches that end a batch
  if (!(MODE == kFused && sp.scheme == kLeapFrog)) {
    st3(P.F, s, F);
    st3(P.T, s, T);
  }
  if (MODE == kForcesOnly) return;

  Kin kin;
  kin.x = me.x;
  kin.v = me.v;
  kin.w = me.w;
  kin.wb = ld3(P.wb, s);
  kin.q = quat{d3{P.q[0][s], P.q[1][s], P.q[2][s]}, P.q[3][s]};
  const double cf = P.cfac[s], I = P.inertia[s];
  const int nv = cv ^ 1;
  if (SHAPES) {
    const d3 I3 = gyro ? d3{P.I3[0][s], P.I3[1][s], P.I3[2][s]} : d3{I, I, I};
    lf_kick_body(kin, gyro, F, T, me.m, cf, I, I3, 0.5 * sp.dt);
    if (MODE == kFused) {
      double depl = lf_move_body(kin, gyro, F, T, me.m, cf, I, I3, 0.5 * sp.dt, sp.dt, B);
      if (depl > me.c) raise_error(st, kMotionError, me.id, 0);
      wrap_and_bin(kin, G, S, st, s, me.id, 0, &VS);
      st3(P.x[cx ^ 1], s, kin.x);
      for (int j = 0; j < 9; ++j) P.B[cx ^ 1][j][s] = B[j];
      P.q[0][s] = kin.q.v.x;
      P.q[1][s] = kin.q.v.y;
      P.q[2][s] = kin.q.v.z;
      P.q[3][s] = kin.q.w;
    }
    st3(P.v[nv], s, kin.v);
    st3(P.w[nv], s, kin.w);
    st3(P.wb, s, kin.wb);
    return;
  }
  sphere_tail<MODE>(P, S, G, sp, st, VS, s, cx, cv, kin, me.m, me.c, me.id, cf, I, F, T);
}

// Per-block sums of the contact statistics (sum and max of the overlaps,
// their count, GJK iterations and evaluations), in a fixed order.
// (ADD: combined with the sums already in out, written by an earlier launch
// over the same particles.)
template <bool ADD = false>
__device__ __forceinline__ void block_stats(const StatAcc& a, double* out) {
  __shared__ double part[4][5];
  double v[5] = {a.sum, a.max, a.cnt, a.git, a.gn};
  for (int off = 32; off > 0; off >>= 1) {
#pragma unroll
    for (int j = 0; j < 5; ++j) {
      const double o = __shfl_down(v[j], off);
      v[j] = j == 1 ? fmax(v[j], o) : v[j] + o;
    }
  }
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (lane == 0)
    for (int j = 0; j < 5; ++j) part[w][j] = v[j];
  __syncthreads();
  if (threadIdx.x == 0) {
    const int nw = (blockDim.x + 63) >> 6;
    for (int j = 0; j < 5; ++j) {
      double r = part[0][j];
      for (int k = 1; k < nw; ++k) r = j == 1 ? fmax(r, part[k][j]) : r + part[k][j];
      if (ADD) r = j == 1 ? fmax(out[j], r) : out[j] + r;
      out[j] = r;
    }
  }
}

// Spheres: one thread per storage index, so that the particle's own state is
// read and written in contiguous runs; its sorted slot locates the cell data.
// Non-spherical bodies: one thread per sorted slot, so that a wave holds
// neighbouring bodies (GJK dominates, and neighbours share branches).
// EXTRAS: contact statistics (sp.stats set) and contact memory models may
// occur, checked at run time.  Spheres have a separate instantiation
// without them, so that the common launch carries none of their cost; the
// GJK kernels, heavy to compile, always check at run time.
template <int MODE, bool MIXED, bool SHAPES, bool EXTRAS>
__global__ __launch_bounds__(256) void k_forces(Particles P, Sorted S, Grid G, const BoxObstacle* __restrict__ obs,
                                                const HookeModel* __restrict__ models, StepParams sp, DevStatus* st,
                                                int* pairs, NbrList L, VerletState VS, NbrList CL) {
  const int t = xcd_tile(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
  if (!EXTRAS) {
    if (t < sp.n) force_one<MODE, MIXED, SHAPES, false>(P, S, G, obs, models, sp, st, pairs, L, VS, CL, t, nullptr);
    return;
  }
  if (SHAPES) {
    // the narrow phase (k_narrow) collected this evaluation's statistics
    if (t < sp.n) force_one<MODE, MIXED, true, true>(P, S, G, obs, models, sp, st, pairs, L, VS, CL, t, nullptr);
    return;
  }
  StatAcc acc;
  if (t < sp.n)
    force_one<MODE, MIXED, SHAPES, true>(P, S, G, obs, models, sp, st, pairs, L, VS, CL, t,
                                         sp.stats ? &acc : nullptr);
  if (sp.stats) block_stats(acc, sp.stats + 5 * (size_t)blockIdx.x);
}

// FP64 spheres without extras: register use capped so that three waves per
// SIMD fit (the uncapped allocation leaves room for two).
template <int MODE>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3))) void k_forces64(
    Particles P, Sorted S, Grid G, const BoxObstacle* __restrict__ obs, const HookeModel* __restrict__ models,
    StepParams sp, DevStatus* st, int* pairs, NbrList L, VerletState VS, NbrList CL) {
  const int t = xcd_tile(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
  if (t < sp.n) force_one<MODE, false, false, false>(P, S, G, obs, models, sp, st, pairs, L, VS, CL, t, nullptr);
}


constexpr int kSphereBlock = 64;
static inline int nblk_f(int n, int b) { return (n + b - 1) / b; }

// Launchers of the force sweep, one explicit instantiation per translation
// unit: spheres only (dem_forces_sph64.hip, dem_forces_sph32.hip) and
// systems with non-spherical bodies (dem_forces_shp64.hip, dem_forces_shp32.hip).
template <int MODE, bool MIXED>
void launch_forces_sph(const Particles& P, const Sorted& S, const Grid& G, const BoxObstacle* obs,
                       const HookeModel* models, const StepParams& sp, DevStatus* st, int* pairs, const NbrList& L,
                       const VerletState& V, const NbrList& CL, hipStream_t stream);
template <int MODE, bool MIXED>
void launch_forces_shp(const Particles& P, const Sorted& S, const Grid& G, const BoxObstacle* obs,
                       const HookeModel* models, const StepParams& sp, DevStatus* st, int* pairs, const NbrList& L,
                       const VerletState& V, const NbrList& CL, hipStream_t stream);

}  // namespace dem
