// Narrow phase of systems with non-spherical bodies, as kernels of their own:
// for every sorted particle, the closest points of its neighbour candidates
// and of the obstacles (Particle::InterAction, Particle.cpp:477-545, with
// RigidBodyWithCrust::ClosestPoint, RigidBodyWithCrust.cpp:205-354, its GJK
// general path and the ContactError recovery of :534-590; SimpleObstacle::
// InterAction, SimpleObstacle.cpp:222-271), recorded as ContactFeat entries
// in the order of the 27-cell sweep.  The force sweep then reads those
// records, so that the GJK state and the integrator never share a kernel's
// registers.  Contact statistics of the evaluation are collected here.
//
//   k_narrow       the listed candidates and the obstacles through ONE
//                  evaluation site (one copy of the GJK state in registers)
//   k_narrow_full  the particles whose candidates overflowed the list (or
//                  every particle without a list): the generic 27-cell sweep
#include <hip/hip_runtime.h>

#include "dem_device.hpp"
#include "dem_forces.hpp"
#include "dem_launch.hpp"

namespace dem {

// Records each contact (overlap < 0) of one particle.
struct FeatureVisitor {
  const ContactFeat& cf;
  int k;
  int c = 0;
  __device__ __forceinline__ void put(int o, int pid, d3 xo, d3 Pc, d3 ov) {
    if (c < cf.cap) {
      const size_t e = (size_t)c * cf.n + k;
      cf.ent[e] = o;
      cf.pid[e] = pid;
      cf.f[0][e] = xo.x;
      cf.f[1][e] = xo.y;
      cf.f[2][e] = xo.z;
      cf.f[3][e] = Pc.x;
      cf.f[4][e] = Pc.y;
      cf.f[5][e] = Pc.z;
      cf.f[6][e] = ov.x;
      cf.f[7][e] = ov.y;
      cf.f[8][e] = ov.z;
    }
    ++c;
  }
  __device__ __forceinline__ void pair(int, int s2i, int pid2, d3 xo, double, d3 Pc, d3 ov, double) {
    put(s2i, pid2, xo, Pc, ov);
  }
  __device__ __forceinline__ void wall(int o, d3 Pc, d3 ov, double) { put(-1 - o, 0, d3{0., 0., 0.}, Pc, ov); }
};

__device__ __forceinline__ bool load_self(const Particles& P, const Sorted& S, const StepParams& sp, int k, Self& me,
                                          int& s) {
  s = S.order[k];
  if ((unsigned)s >= (unsigned)sp.n) return false;
  const double* pr = S.posr + 4 * (size_t)k;
  me.x = d3{pr[0], pr[1], pr[2]};
  me.r = pr[3];
  me.v = me.w = d3{0., 0., 0.};
  me.m = me.req = 0.;
  me.mat = 0;
  me.id = P.id[s];
  me.c = P.crust[s];
  return true;
}

// The listed candidates (count <= cap) and the obstacles of sorted slot k,
// in the order of for_each_listed, each evaluated at the one site below.
// Nothing but the loop state stays live across that site: both bodies are
// rebuilt from memory at every pass, and a pair whose crust-shrunk bodies
// intersect comes round once more for the recovery (attempt 1).  A particle
// without a usable list is marked (count -1) for k_narrow_full.
__device__ __forceinline__ void narrow_one(const Particles& P, const Sorted& S, const Grid& G,
                                           const BoxObstacle* __restrict__ obs, const StepParams& sp, DevStatus* st,
                                           const NbrList& L, int k, StatAcc* acc) {
  const int cx = sp.cur & 1;
  Self me;
  int s;
  if (!load_self(P, S, sp, k, me, s)) {
    sp.cf.count[k] = 0;
    return;
  }
  const int nc = L.count ? L.count[k] : 0;
  if (!L.count || nc > L.cap) {
    sp.cf.count[k] = -1;
    return;
  }
  const int self_type = P.shape[s] < 0 ? kShapeSphere : P.shapes[P.shape[s]].type;
  FeatureVisitor vis{sp.cf, k};
  int attempt = 0;
  for (int j = 0; j < nc + sp.nobs; j += attempt ? 0 : 1) {
    const int att = attempt;
    attempt = 0;
    d3 Pc, ov, xo{0., 0., 0.};
    double d = 0.;
    int nit = 1, res = 0, partner, pid2, mult;
    ConvexGeo og;
    bool convex = false;
    if (j < nc) {
      // visit_candidate: periodic image, conservative reject
      const unsigned e = L.ent[(size_t)j * L.n + k];
      const int k2 = (int)(e & kNbrSlotMask), code = (int)(e >> kNbrSlotBits);
      const int ix = code % 3 - 1, iy = (code / 3) % 3 - 1, iz = code / 9 - 1;
      const double* q2 = S.posr + 4 * (size_t)k2;
      xo = d3{ix ? q2[0] + image_shift(G, 0, ix) : q2[0], iy ? q2[1] + image_shift(G, 1, iy) : q2[1],
              iz ? q2[2] + image_shift(G, 2, iz) : q2[2]};
      const double ro = q2[3];
      const d3 AB = xo - me.x;
      const double s2 = AB.x * AB.x + AB.y * AB.y + AB.z * AB.z;
      const double R = me.r + ro;
      if (s2 > R * R * 1.000000000001) continue;
      const int s2i = S.order[k2];
      if ((unsigned)s2i >= (unsigned)sp.n) continue;
      partner = s2i;
      pid2 = S.sid[k2];
      const int sh[3] = {ix, iy, iz};
      mult = me.id < pid2 ? image_evaluations(G, me.x, d3{q2[0], q2[1], q2[2]}, sh) : 0;
      const bool osph = P.shape[s2i] < 0 || P.shapes[P.shape[s2i]].type == kShapeSphere;
      if (self_type == kShapeSphere && osph) {
        const double co = P.crust[s2i];
        res = sphere_sphere(me.x, me.r, me.c, AB, sqrt(s2), ro, co, Pc, ov, d);
        if (res < 0) res = recover_ss(me.x, me.r, me.c, xo, ro, co, Pc, ov, d) ? 2 : -1;
      } else {
        og = convex_of(P, s2i, xo, cx);
        convex = true;
      }
    } else {
      // SimpleObstacle::InterAction: bounding-box pre-test
      const int o = j - nc;
      const BoxObstacle& B = obs[o];
      const bool probable = fabs(B.aabb_c[0] - me.x.x) <= B.aabb_e[0] + me.r &&
                            fabs(B.aabb_c[1] - me.x.y) <= B.aabb_e[1] + me.r &&
                            fabs(B.aabb_c[2] - me.x.z) <= B.aabb_e[2] + me.r;
      if (!probable) continue;
      partner = -1 - o;
      pid2 = B.id;
      mult = 1;
      if (self_type == kShapeSphere) {
        res = sphere_box(B, me.x, me.r, me.c, Pc, ov, d);
        if (res < 0) res = recover_sb(B, me.x, me.r, me.c, Pc, ov, d) ? 2 : -1;
      } else {
        og = convex_of_obstacle(B);
        convex = true;
      }
    }
    if (convex) {
      const ConvexGeo mg = convex_of(P, s, me.x, cx);
      // closest_convex: sphere-box closed forms (a closed form that finds
      // the crusts overlapping recovers through GJK), otherwise GJK
      const bool sb = mg.type == kShapeSphere && og.type == kShapeBox;
      const bool bs = mg.type == kShapeBox && og.type == kShapeSphere;
      if ((sb || bs) && att == 0) {
        res = sb ? sphere_box_particle(og, mg, true, Pc, ov, d) : sphere_box_particle(mg, og, false, Pc, ov, d);
        if (res < 0) {
          attempt = 1;
          continue;
        }
      } else {
        res = gjk_attempt(mg, og, att, Pc, ov, d, nit, acc == nullptr);
        if (res < 0 && att == 0) {
          attempt = 1;
          continue;
        }
      }
    }
    if (res == 0) continue;
    if (res < 0) {
      raise_error(st, kContactError, me.id, pid2);
      continue;
    }
    if (res == 2) atomicAdd(&st->n_recovered, 1);
    if (acc && mult) acc->add(d, nit, mult);
    if (!(d < 0.)) continue;
    vis.put(partner, pid2, xo, Pc, ov);
  }
  sp.cf.count[k] = vis.c;
  if (vis.c > sp.cf.cap) raise_error(st, kContactOverflow, me.id, vis.c);
}

// The generic 27-cell sweep for the particles k_narrow marked.
__device__ __forceinline__ void narrow_full_one(const Particles& P, const Sorted& S, const Grid& G,
                                                const BoxObstacle* __restrict__ obs, const StepParams& sp,
                                                DevStatus* st, int k, StatAcc* acc) {
  const int cx = sp.cur & 1;
  Self me;
  int s;
  if (!load_self(P, S, sp, k, me, s)) return;
  me.acc = acc;
  const ConvexGeo mg = convex_of(P, s, me.x, cx);
  FeatureVisitor vis{sp.cf, k};
  for_each_contact<FeatureVisitor, true>(P, S, G, obs, sp.nobs, st, 0, 0, sp.n, k, me, vis, &mg, cx);
  sp.cf.count[k] = vis.c;
  if (vis.c > sp.cf.cap) raise_error(st, kContactOverflow, me.id, vis.c);
}

// One thread per sorted slot; with statistics, 256-thread workgroups whose
// partial sums k_stats_fold adds up.
__global__ __launch_bounds__(256) void k_narrow(Particles P, Sorted S, Grid G, const BoxObstacle* __restrict__ obs,
                                                StepParams sp, DevStatus* st, NbrList L) {
  const int k = xcd_tile(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
  if (!sp.stats) {
    if (k < sp.n) narrow_one(P, S, G, obs, sp, st, L, k, nullptr);
    return;
  }
  StatAcc acc;
  if (k < sp.n) narrow_one(P, S, G, obs, sp, st, L, k, &acc);
  block_stats(acc, sp.stats + 5 * (size_t)blockIdx.x);
}

__global__ __launch_bounds__(256) void k_narrow_full(Particles P, Sorted S, Grid G,
                                                     const BoxObstacle* __restrict__ obs, StepParams sp,
                                                     DevStatus* st) {
  const int k = xcd_tile(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
  const bool mine = k < sp.n && sp.cf.count[k] < 0;
  if (!sp.stats) {
    if (mine) narrow_full_one(P, S, G, obs, sp, st, k, nullptr);
    return;
  }
  StatAcc acc;
  if (mine) narrow_full_one(P, S, G, obs, sp, st, k, &acc);
  block_stats<true>(acc, sp.stats + 5 * (size_t)blockIdx.x);
}

hipError_t launch_narrow(const Particles& P, const Sorted& S, const Grid& G, const BoxObstacle* obs,
                         const StepParams& sp, DevStatus* st, const NbrList& L, hipStream_t stream) {
  if (sp.n == 0) return hipSuccess;
  const int bs = sp.stats ? 256 : 64;
  const dim3 grid((sp.n + bs - 1) / bs);
  hipLaunchKernelGGL(k_narrow, grid, dim3(bs), 0, stream, P, S, G, obs, sp, st, L);
  hipLaunchKernelGGL(k_narrow_full, grid, dim3(bs), 0, stream, P, S, G, obs, sp, st);
  return hipGetLastError();
}

}  // namespace dem
