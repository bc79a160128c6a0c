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
//   k_items        one thread per sorted particle: the work list of the
//                  step -- one item per listed candidate the particle
//                  evaluates and per obstacle whose bounding box it meets
//                  -- appended with one atomic per wave; the obstacle
//                  record slots it may fill are marked kNoFeat
//   k_narrow_items a grid of one-wave workgroups over the work list, one
//                  evaluation per lane, so that the GJK work of a dense bed
//                  occupies as many lanes as it has pairs instead of one
//                  lane per body; a contact's record goes to the item's
//                  slot (candidate j: slot j; obstacle o: slot nc + o)
//   k_narrow       with contact statistics: one thread per particle over
//                  all its candidates and obstacles, records packed;
//                  through ONE evaluation site (one copy of the GJK state
//                  in registers)
//   k_narrow_full  the particles whose candidates overflowed the list (or
//                  every particle without a list): the generic 27-cell sweep
//
// The force sweep reads slots 0 .. count-1 in order and skips kNoFeat, so
// both layouts give the same contacts in the same order.
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
    put_at(c, o, pid, xo, Pc, ov);
    ++c;
  }
  __device__ __forceinline__ void put_at(int slot, int o, int pid, d3 xo, d3 Pc, d3 ov) {
    if (slot < cf.cap) {
      const size_t e = (size_t)slot * cf.n + k;
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

// A pair without periodic shift whose evaluation is exactly symmetric under
// the exchange of the bodies (every path but the sphere-sphere closed form,
// which places the point from body A's centre) is evaluated once, by its
// lower id, which also writes the partner's record -- the mirror image:
// same point, opposite overlap vector -- unless the partner has no usable
// list (k_narrow_full then covers it).  The partner's record slot holds
// kNoFeat from k_detect until then.
__device__ __forceinline__ bool mirrored_pair(const NbrList& L, int code, bool both_spheres, int k2) {
  return code == kNbrNoShift && !both_spheres && L.count[k2] <= L.cap;
}

// ITEM: the one evaluation `item` of sorted slot k (a listed candidate if
// item < nc, else obstacle item - nc), its record at slot `item`.
// Otherwise: every candidate and obstacle of k, in the order of
// for_each_listed, records packed and counted.  Each evaluation goes
// through the one site below; nothing but the loop state stays live across
// it: both bodies are rebuilt from memory at every pass, and a pair whose
// crust-shrunk bodies intersect comes round once more for the recovery
// (attempt 1).  A particle without a usable list is marked (count -1) for
// k_narrow_full.
template <bool ITEM>
__device__ __forceinline__ void narrow_one(const Particles& P, const Sorted& S, const Grid& G,
                                           const BoxObstacle* __restrict__ obs, const StepParams& sp, DevStatus* st,
                                           const NbrList& L, int k, int item, StatAcc* acc) {
  const int cx = sp.cur & 1;
  Self me;
  int s;
  if (!load_self(P, S, sp, k, me, s)) {
    if (!ITEM) sp.cf.count[k] = 0;
    return;
  }
  const int nc = L.count ? L.count[k] : 0;
  if (!L.count || nc > L.cap) {
    if (!ITEM) sp.cf.count[k] = -1;
    return;
  }
  const int self_type = P.shape[s] < 0 ? kShapeSphere : P.shapes[P.shape[s]].type;
  FeatureVisitor vis{sp.cf, k};
  int attempt = 0;
  const int nitems = ITEM ? 1 : nc + sp.nobs;
  for (int it = 0; it < nitems; it += attempt ? 0 : 1) {
    const int j = ITEM ? item : it;
    const int att = attempt;
    attempt = 0;
    d3 Pc, ov, xo{0., 0., 0.};
    double d = 0.;
    int nit = 1, res = 0, partner, pid2, mult;
    int k2 = -1;          // the partner's sorted slot (candidates)
    bool mirror = false;  // this thread also writes the partner's record
    ConvexGeo og;
    bool convex = false;
    if (j < nc) {
      // visit_candidate: periodic image, conservative reject
      const unsigned e = L.ent[(size_t)j * L.n + k];
      k2 = (int)(e & kNbrSlotMask);
      const int code = (int)(e >> kNbrSlotBits);
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
      if (ITEM && mirrored_pair(L, code, self_type == kShapeSphere && osph, k2)) {
        if (me.id > pid2) continue;  // the partner's item (k_items lists none here)
        mirror = true;
      }
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
    if (res == 2) atomicAdd(&st->n_recovered, mirror ? 2 : 1);  // counted on each body's side
    if (acc && mult) acc->add(d, nit, mult);
    if (!(d < 0.)) continue;
    if (ITEM) {
      vis.put_at(j, partner, pid2, xo, Pc, ov);
      if (mirror) {
        // this particle's slot in the partner's list (the lists are
        // symmetric: the candidate test is exact under the exchange)
        const int nc2 = L.count[k2];
        const unsigned want = nbr_entry(k, 0, 0, 0);
        int jm = -1;
        for (int i0 = 0; i0 < nc2 && jm < 0; i0 += 16) {
          unsigned hit = 0u;
#pragma unroll
          for (int u = 0; u < 16; ++u)
            if (i0 + u < nc2 && L.ent[(size_t)(i0 + u) * L.n + k2] == want) hit |= 1u << u;
          if (hit) jm = i0 + __builtin_ctz(hit);
        }
        if (jm < 0) {
          raise_error(st, kContactOverflow, me.id, pid2);
        } else {
          FeatureVisitor pv{sp.cf, k2};
          pv.put_at(jm, s, me.id, me.x, Pc, -ov);
        }
      }
    } else {
      vis.put(partner, pid2, xo, Pc, ov);
    }
  }
  if (ITEM) return;
  sp.cf.count[k] = vis.c;
  if (vis.c > sp.cf.cap) raise_error(st, kContactOverflow, me.id, vis.c);
}

// The work list of sorted slot k (see k_items); returns its number of items
// and, with out != nullptr, writes them from out[0].  With keep != nullptr
// (out == nullptr) it also returns which candidates (keep[0], bit j) and
// obstacles (keep[1], bit o) give items and the last obstacle met (keep[2];
// -2: no body in the slot), for write_items.
__device__ __forceinline__ int body_items(const Particles& P, const Sorted& S, const BoxObstacle* __restrict__ obs,
                                          const StepParams& sp, const NbrList& L, int k,
                                          unsigned long long* __restrict__ out, long long* keep = nullptr) {
  const int s = S.order[k];
  if ((unsigned)s >= (unsigned)sp.n) {
    if (out) sp.cf.count[k] = 0;
    if (keep) keep[2] = -2;
    return 0;
  }
  const int nc = L.count[k];
  if (nc > L.cap) return 0;  // k_narrow_full
  const double* pr = S.posr + 4 * (size_t)k;
  const double x[3] = {pr[0], pr[1], pr[2]}, r = pr[3];
  const int id = P.id[s];
  const bool ssph = P.shape[s] < 0 || P.shapes[P.shape[s]].type == kShapeSphere;
  const unsigned long long key = (unsigned long long)(unsigned)k << 32;
  int c = 0;
  for (int j = 0; j < nc; ++j) {
    const unsigned e = L.ent[(size_t)j * L.n + k];
    const int k2 = (int)(e & kNbrSlotMask), code = (int)(e >> kNbrSlotBits);
    if (code == kNbrNoShift) {
      const int s2i = S.order[k2];
      if ((unsigned)s2i < (unsigned)sp.n) {
        const bool osph = P.shape[s2i] < 0 || P.shapes[P.shape[s2i]].type == kShapeSphere;
        if (mirrored_pair(L, code, ssph && osph, k2) && id > S.sid[k2]) continue;
      }
    }
    if (out) out[c] = key | (unsigned)j;
    if (keep && j < 64) keep[0] |= 1ll << j;
    ++c;
  }
  int last = -1;
  for (int o = 0; o < sp.nobs; ++o) {
    const BoxObstacle& B = obs[o];
    if (fabs(B.aabb_c[0] - x[0]) <= B.aabb_e[0] + r && fabs(B.aabb_c[1] - x[1]) <= B.aabb_e[1] + r &&
        fabs(B.aabb_c[2] - x[2]) <= B.aabb_e[2] + r) {
      if (out) out[c] = key | (unsigned)(nc + o);
      if (keep && o < 64) keep[1] |= 1ll << o;
      ++c;
      last = o;
    }
  }
  if (keep) keep[2] = last;
  if (out) {
    // every obstacle slot the force sweep will read, met or not, starts empty
    for (int o = 0; o <= last; ++o) sp.cf.ent[(size_t)(nc + o) * sp.cf.n + k] = kNoFeat;
    sp.cf.count[k] = nc + last + 1;
  }
  return c;
}

// The items body_items found for slot k, written from the masks it kept
// (no second pass over the candidate list), in the same order.
__device__ __forceinline__ void write_items(const StepParams& sp, int k, int nc, const long long* keep,
                                            unsigned long long* __restrict__ out) {
  if (keep[2] == -2) {
    sp.cf.count[k] = 0;
    return;
  }
  const unsigned long long key = (unsigned long long)(unsigned)k << 32;
  int c = 0;
  for (unsigned long long m = (unsigned long long)keep[0]; m; m &= m - 1)
    out[c++] = key | (unsigned)__builtin_ctzll(m);
  for (unsigned long long m = (unsigned long long)keep[1]; m; m &= m - 1)
    out[c++] = key | (unsigned)(nc + __builtin_ctzll(m));
  const int last = (int)keep[2];
  for (int o = 0; o <= last; ++o) sp.cf.ent[(size_t)(nc + o) * sp.cf.n + k] = kNoFeat;
  sp.cf.count[k] = nc + last + 1;
}

// One thread per sorted slot: counts its items, one atomic per wave reserves
// the wave's run of the list, then each lane writes its items (from the
// masks of the count pass when the candidates and obstacles fit 64 bits).
// Particles without a usable list get count -1 (k_narrow_full).
__global__ __launch_bounds__(256) void k_items(Particles P, Sorted S, const BoxObstacle* __restrict__ obs,
                                               StepParams sp, NbrList L, const DevStatus* st) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  const bool mine = k < sp.n && !st->code;  // after an error: an empty list
  if (mine && L.count[k] > L.cap) sp.cf.count[k] = -1;
  long long keep[3] = {0, 0, -1};
  const int c = mine ? body_items(P, S, obs, sp, L, k, nullptr, keep) : 0;
  // inclusive scan of the counts over the wave
  const int lane = threadIdx.x & 63;
  int incl = c;
  for (int off = 1; off < 64; off <<= 1) {
    const int v = __shfl_up(incl, off);
    if (lane >= off) incl += v;
  }
  const int total = __shfl(incl, 63);
  int base = 0;
  if (lane == 63 && total > 0) base = atomicAdd(sp.cf.nitems, total);
  base = __shfl(base, 63);
  if (mine && L.count[k] <= L.cap) {
    unsigned long long* out = sp.cf.items + base + (incl - c);
    if (L.cap <= 64 && sp.nobs <= 64) write_items(sp, k, L.count[k], keep, out);
    else body_items(P, S, obs, sp, L, k, out);
  }
}

// The work list, one item per lane at a time (a grid of one-wave
// workgroups that strides over it).
__global__ __launch_bounds__(64) void k_narrow_items(Particles P, Sorted S, Grid G, const BoxObstacle* __restrict__ obs,
                                                     StepParams sp, DevStatus* st, NbrList L) {
  const int total = *sp.cf.nitems;
  for (int i = blockIdx.x * 64 + threadIdx.x; i < total; i += gridDim.x * 64) {
    const unsigned long long it = sp.cf.items[i];
    narrow_one<true>(P, S, G, obs, sp, st, L, (int)(it >> 32), (int)(unsigned)it, nullptr);
  }
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

// One thread per sorted slot, records packed: with statistics (256-thread
// workgroups whose partial sums k_stats_fold adds up), and for lists too
// long for the slot layout.
__global__ __launch_bounds__(256) void k_narrow(Particles P, Sorted S, Grid G, const BoxObstacle* __restrict__ obs,
                                                StepParams sp, DevStatus* st, NbrList L) {
  const int k = xcd_tile(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
  const bool mine = k < sp.n && !st->code;  // an error ended the run
  if (!sp.stats) {
    if (mine) narrow_one<false>(P, S, G, obs, sp, st, L, k, 0, nullptr);
    return;
  }
  StatAcc acc;
  if (mine) narrow_one<false>(P, S, G, obs, sp, st, L, k, 0, &acc);
  block_stats(acc, sp.stats + 5 * (size_t)blockIdx.x);
}

__global__ __launch_bounds__(256) void k_narrow_full(Particles P, Sorted S, Grid G,
                                                     const BoxObstacle* __restrict__ obs, StepParams sp,
                                                     DevStatus* st) {
  const int k = xcd_tile(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
  const bool mine = k < sp.n && sp.cf.count[k] < 0 && !st->code;
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
  // packed records where slots would not fit the obstacles' records
  if (sp.stats || !L.count || L.cap + sp.nobs > sp.cf.cap) {
    const dim3 grid((sp.n + 255) / 256);
    hipLaunchKernelGGL(k_narrow, grid, dim3(256), 0, stream, P, S, G, obs, sp, st, L);
    hipLaunchKernelGGL(k_narrow_full, grid, dim3(256), 0, stream, P, S, G, obs, sp, st);
    return hipGetLastError();
  }
  hipError_t e = launch_clear(sp.cf.nitems, 1, stream);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_items, dim3((sp.n + 255) / 256), dim3(256), 0, stream, P, S, obs, sp, L, st);
  hipLaunchKernelGGL(k_narrow_items, dim3(sp.cf.item_blocks), dim3(64), 0, stream, P, S, G, obs, sp, st, L);
  hipLaunchKernelGGL(k_narrow_full, dim3((sp.n + 63) / 64), dim3(64), 0, stream, P, S, G, obs, sp, st);
  return hipGetLastError();
}

}  // namespace dem
