// HIP kernels of the Grains3D DEM time step for gfx950 (MI355X).
//
// Per step (SecondOrderLeapFrog, kick-drift-kick, Grains.cpp:344-364):
//   k_move / fused tail  half kick + drift + quaternion update, periodic role
//                        swap, cell key, atomic per-cell arrival slot
//   k_scan_*             exclusive prefix of the per-cell counts -> start[]
//   k_scatter            storage index to (cell, arrival slot)
//   k_canon              per cell: order by particle id, gather (x,y,z,r) into
//                        the sorted tile array, reset the counts
//   k_detect             one thread per sorted particle: 27-cell sweep with
//                        minimum-image shifts, neighbour candidates -> list
//   k_forces<MODE>       one thread per particle (storage order; sorted order
//                        with non-spherical bodies): the listed (or, with no
//                        list, the 27-cell) candidates with
//                        minimum-image shifts, sphere-sphere / sphere-box
//                        closest points, Hooke-dashpot-Coulomb force and
//                        torque gathered in registers (no atomics), then the
//                        second half kick, optionally fused with the next
//                        step's first half kick and drift.
//
// FP64 arithmetic follows the reference expression order and the library is
// built with -ffp-contract=off; the mixed kernels evaluate the force model in
// FP32 (contact decisions stay FP64 so the contact list is identical).
#include <hip/hip_runtime.h>
#include <algorithm>
#include <cstdint>

#include "dem_device.hpp"
#include "dem_forces.hpp"
#include "dem_launch.hpp"

namespace dem {

// ---------------------------------------------------------------------------
// First half kick + drift over storage indices (in place): Grains.cpp:353
// moveParticlesAndObstacles(dt/2, dt, dt) -> AllComponents::Move.
__global__ __launch_bounds__(256) void k_move(Particles P, Sorted S, Grid G, StepParams sp, DevStatus* st,
                                              double dtv, double dtx, VerletState V) {
  int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= sp.n) return;
  const int cx = sp.cur & 1, cv = (sp.cur >> 1) & 1;
  Kin k = load_kin(P, s, cx, cv);
  d3 F = ld3(P.F, s), T = ld3(P.T, s);
  double mass = P.mass[s], cf = P.cfac[s], I = P.inertia[s];
  double depl;
  if (P.shape) {
    // system with non-spherical particles (LeapFrog only): basis in place
    double B[9];
    for (int j = 0; j < 9; ++j) B[j] = P.B[cx][j][s];
    const int sh = P.shape[s];
    const bool gyro = sh >= 0 && P.shapes[sh].type != kShapeSphere;
    const d3 I3 = gyro ? d3{P.I3[0][s], P.I3[1][s], P.I3[2][s]} : d3{I, I, I};
    depl = lf_move_body(k, gyro, F, T, mass, cf, I, I3, dtv, dtx, B);
    for (int j = 0; j < 9; ++j) P.B[cx][j][s] = B[j];
  } else if (sp.scheme == kLeapFrog) {
    depl = lf_move(k, F, T, mass, cf, I, dtv, dtx);
  } else {
    AB2State ab;
    if (sp.scheme == kAdamsBashforth2) {
      ab.v_nm2 = ld3(P.v_nm2, s);
      ab.dUdt_nm2 = ld3(P.dUdt_nm2, s);
      ab.wb_nm2 = ld3(P.wb_nm2, s);
      ab.dWdt_nm2 = ld3(P.dWdt_nm2, s);
    }
    depl = other_move(sp.scheme, sp.first_step != 0, k, ab, F, T, mass, cf, I, dtv, dtx);
    if (sp.scheme == kAdamsBashforth2) {
      st3(P.v_nm2, s, ab.v_nm2);
      st3(P.dUdt_nm2, s, ab.dUdt_nm2);
      st3(P.wb_nm2, s, ab.wb_nm2);
      st3(P.dWdt_nm2, s, ab.dWdt_nm2);
    }
  }
  int pid = P.id[s];
  if (depl > P.crust[s]) raise_error(st, kMotionError, pid, 0);  // Particle.cpp:444-456
  wrap_and_bin(k, G, S, st, s, pid, 0, &V);
  st3(P.x[cx], s, k.x);
  st3(P.v[cv], s, k.v);
  st3(P.w[cv], s, k.w);
  st3(P.wb, s, k.wb);
  P.q[0][s] = k.q.v.x;
  P.q[1][s] = k.q.v.y;
  P.q[2][s] = k.q.v.z;
  P.q[3][s] = k.q.w;
}

// Cell key and arrival slot of the current positions, without motion (used
// after initialisation and after a state upload).  Positions outside the
// periodic box are brought back first.
__global__ __launch_bounds__(256) void k_key(Particles P, Sorted S, Grid G, StepParams sp, DevStatus* st,
                                             VerletState V) {
  int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= sp.n) return;
  const int cx = sp.cur & 1, cv = (sp.cur >> 1) & 1;
  Kin k = load_kin(P, s, cx, cv);
  d3 x0 = k.x;
  // with a Verlet list the periodic wrap only (the list is rebuilt next)
  wrap_and_bin(k, G, S, st, s, P.id[s], 0, &V);
  if (k.x.x != x0.x || k.x.y != x0.y || k.x.z != x0.z) {
    st3(P.x[cx], s, k.x);
    st3(P.wb, s, k.wb);
  }
}

// ---------------------------------------------------------------------------
// Exclusive scan of the per-cell counts.  Tiles of 256 threads x 8 cells.
constexpr int kScanThreads = 256;
constexpr int kScanItems = 8;
constexpr int kScanTile = kScanThreads * kScanItems;

__device__ __forceinline__ int block_exclusive_scan(int v, int* lds, int& total) {
  // Hillis-Steele over 256 entries in LDS
  int t = threadIdx.x;
  lds[t] = v;
  __syncthreads();
  for (int off = 1; off < kScanThreads; off <<= 1) {
    int add = t >= off ? lds[t - off] : 0;
    __syncthreads();
    lds[t] += add;
    __syncthreads();
  }
  int incl = lds[t];
  total = lds[kScanThreads - 1];
  __syncthreads();
  return incl - v;
}

__global__ __launch_bounds__(kScanThreads) void k_scan_reduce(const int* __restrict__ count, int ncells,
                                                              int* __restrict__ block_sums, const int* gate,
                                                              int stamp) {
  if (gate && *gate != stamp) return;  // Verlet list still valid: nothing to rebuild
  __shared__ int lds[kScanThreads];
  int base = blockIdx.x * kScanTile + threadIdx.x * kScanItems;
  int sum = 0;
#pragma unroll
  for (int i = 0; i < kScanItems; ++i) sum += (base + i < ncells) ? count[base + i] : 0;
  int total;
  block_exclusive_scan(sum, lds, total);
  if (threadIdx.x == 0) block_sums[blockIdx.x] = total;
}

__global__ __launch_bounds__(kScanThreads) void k_scan_top(int* __restrict__ block_sums, int nblocks, const int* gate,
                                                           int stamp) {
  if (gate && *gate != stamp) return;  // Verlet list still valid: nothing to rebuild
  __shared__ int lds[kScanThreads];
  int carry = 0;
  for (int b0 = 0; b0 < nblocks; b0 += kScanThreads) {
    int i = b0 + threadIdx.x;
    int v = i < nblocks ? block_sums[i] : 0;
    int total;
    int ex = block_exclusive_scan(v, lds, total);
    if (i < nblocks) block_sums[i] = ex + carry;
    carry += total;
  }
}

// FOLD: block_sums holds the tiles' totals (no k_scan_top launch) and each
// block adds those before it itself -- for grids of at most kScanFoldTiles
// tiles, where the sums are cheaper than a launch.
constexpr int kScanFoldTiles = 4096;

template <bool FOLD>
__global__ __launch_bounds__(kScanThreads) void k_scan_down(const int* __restrict__ count, int ncells,
                                                            const int* __restrict__ block_sums,
                                                            int* __restrict__ start, int n, const int* gate,
                                                            int stamp) {
  if (gate && *gate != stamp) return;  // Verlet list still valid: nothing to rebuild
  __shared__ int lds[kScanThreads];
  int offset;
  if (FOLD) {
    int part = 0;
    for (int b = threadIdx.x; b < (int)blockIdx.x; b += kScanThreads) part += block_sums[b];
    block_exclusive_scan(part, lds, offset);
  } else {
    offset = block_sums[blockIdx.x];
  }
  int base = blockIdx.x * kScanTile + threadIdx.x * kScanItems;
  int vals[kScanItems];
  int sum = 0;
#pragma unroll
  for (int i = 0; i < kScanItems; ++i) {
    vals[i] = (base + i < ncells) ? count[base + i] : 0;
    sum += vals[i];
  }
  int total;
  int run = block_exclusive_scan(sum, lds, total) + offset;
#pragma unroll
  for (int i = 0; i < kScanItems; ++i) {
    if (base + i < ncells) start[base + i] = run;
    run += vals[i];
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) start[ncells] = n;
}

// Storage index -> (cell start + arrival slot)
__global__ __launch_bounds__(256) void k_scatter(Sorted S, int n, const int* gate, int stamp) {
  if (gate && *gate != stamp) return;  // Verlet list still valid: nothing to rebuild
  for (int s = blockIdx.x * blockDim.x + threadIdx.x; s < n; s += gridDim.x * blockDim.x) {
    int dst = S.start[S.key[s]] + S.slot[s];
    if ((unsigned)dst < (unsigned)n) S.tmp_order[dst] = s;
  }
}

// One thread per cell: order the cell's particles by id (deterministic, and
// independent of the atomic arrival order), gather the sorted tile arrays and
// reset the count for the next step.
__global__ __launch_bounds__(256) void k_canon(Particles P, Sorted S, int ncells, int cx, int n, const int* gate,
                                               int stamp) {
  if (gate && *gate != stamp) return;  // Verlet list still valid: nothing to rebuild
  for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < ncells; c += gridDim.x * blockDim.x) {
    S.count[c] = 0;
    canon_cell(P, S, min(S.start[c], n), min(S.start[c + 1], n), cx, 0, n);
  }
}

// ---------------------------------------------------------------------------
// Broad phase of one step (LinkedCell::ComputeForces, LinkedCell.cpp:1019-1103):
// every neighbour candidate of each sorted particle that passes the
// conservative distance reject of the narrow phase, in the order the force
// sweep visits them, written to the candidate list.  Holding only the
// particle's position and the nine row ranges, the kernel keeps few
// registers live, so many waves per SIMD overlap its dependent loads; the
// force sweep then reads a short list instead of 27 cells.
// mark (systems with non-spherical bodies): each listed slot's contact
// record starts as kNoFeat, so that the narrow phase writes only contacts.
// After an error raised on the device (which ends the run) the contact
// searches return at once, so that the rest of a batch of steps costs
// little whatever the state has become.
__global__ __launch_bounds__(256) void k_detect(Sorted S, Grid G, int n, NbrList L, int* __restrict__ mark,
                                                const DevStatus* st) {
  const int k = xcd_tile(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
  if (k >= n || st->code) return;
  const double* pr = S.posr + 4 * (size_t)k;
  const d3 x{pr[0], pr[1], pr[2]};
  const double r = pr[3];
  int c = 0;
  // one candidate at a time: at this register count eight waves per SIMD
  // overlap the loads better than wider batches of loads per thread do
  for_each_candidate(S, G, 0, 0, n, x, [&](int k2, int ix, int iy, int iz) {
    if (k2 == k) return;
    const double* q2 = S.posr + 4 * (size_t)k2;
    d3 xo{ix ? q2[0] + image_shift(G, 0, ix) : q2[0], iy ? q2[1] + image_shift(G, 1, iy) : q2[1],
          iz ? q2[2] + image_shift(G, 2, iz) : q2[2]};
    d3 AB = xo - x;
    double s2 = AB.x * AB.x + AB.y * AB.y + AB.z * AB.z;
    double R = r + q2[3];
    if (s2 > R * R * 1.000000000001) return;
    if (c < L.cap) {
      L.ent[(size_t)c * n + k] = nbr_entry(k2, ix, iy, iz);
      if (mark) mark[(size_t)c * n + k] = kNoFeat;
    }
    ++c;
  });
  L.count[k] = c;
}

// ---------------------------------------------------------------------------
// Force sweep.  MODE: 0 forces only; 1 forces + second half kick; 2 forces +
// second half kick + next step's Move (LeapFrog) or forces + Move(dt, dt)
// (the other schemes).
// Running totals of the contact statistics after a force launch of nb
// blocks: tot = {overlap sum, overlap max, time of the max, contacts, GJK
// iterations, evaluations with iterations}; the maximum keeps the first
// time it was reached (AppCollision.cpp:174-178).
__global__ __launch_bounds__(256) void k_stats_fold(const double* __restrict__ part, int nb, double* tot,
                                                    double time) {
  StatAcc a;
  for (int i = threadIdx.x; i < nb; i += blockDim.x) {
    const double* p = part + 5 * (size_t)i;
    a.sum += p[0];
    a.max = fmax(a.max, p[1]);
    a.cnt += p[2];
    a.git += p[3];
    a.gn += p[4];
  }
  __shared__ double r[5];
  block_stats(a, r);
  __syncthreads();
  if (threadIdx.x == 0) {
    tot[0] += r[0];
    if (r[1] > tot[1]) {
      tot[1] = r[1];
      tot[2] = time;
    }
    tot[3] += r[2];
    tot[4] += r[3];
    tot[5] += r[4];
  }
}

hipError_t launch_stats_fold(const double* part, int n, double* tot, double time, hipStream_t stream) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_stats_fold, dim3(1), dim3(256), 0, stream, part, (n + 255) / 256, tot, time);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Contacts of a step with the Verlet list: the listed candidates of each
// sphere that pass the narrow phase's conservative distance reject at the
// current positions, in the list's order.  Few registers, so that many
// waves overlap the dependent loads of the list; the force sweep then
// visits only these pairs (same pairs, same order, same arithmetic).
__global__ __launch_bounds__(256) void k_vfilter(Particles P, Sorted S, Grid G, VerletState VS, int n, int cx,
                                                 NbrList L, NbrList CL, DevStatus* st) {
  // the list is current (rebuilt if it was stale): clear the stale flag for
  // this step's drift
  if (blockIdx.x == 0 && threadIdx.x == 0) st->rebuild = 0;
  const int s = xcd_tile(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
  if (s >= n || st->code) return;
  const d3 x = ld3(P.x[cx], s);
  const double r = P.rad[s];
  int c2 = 0;
  auto test = [&](int s2i) {
    const d3 xo = nearest_image(G, d3{P.x[cx][0][s2i], P.x[cx][1][s2i], P.x[cx][2][s2i]}, x);
    const d3 AB = xo - x;
    const double s2 = AB.x * AB.x + AB.y * AB.y + AB.z * AB.z;
    const double R = r + P.rad[s2i];
    if (s2 > R * R * 1.000000000001) return;
    if (c2 < CL.cap) CL.ent[(size_t)c2 * n + s] = (unsigned)s2i;
    ++c2;
  };
  const int c = L.count[s];
  if (c <= L.cap) {
    for (int i = 0; i < c; ++i) {
      const unsigned s2i = L.ent[(size_t)i * L.n + s];
      if (s2i < (unsigned)n) test((int)s2i);
    }
  } else {
    const d3 xb{VS.xb[0][s], VS.xb[1][s], VS.xb[2][s]};
    for_each_candidate(S, VS.Gv, 0, 0, n, xb, [&](int k2, int, int, int) {
      const int s2i = S.order[k2];
      if (s2i != s && (unsigned)s2i < (unsigned)n) test(s2i);
    });
  }
  CL.count[s] = c2;
}

// ---------------------------------------------------------------------------
// Host-side launchers
static inline int nblk(int n, int b) { return (n + b - 1) / b; }

hipError_t launch_move(const Particles& P, const Sorted& S, const Grid& G, const StepParams& sp, DevStatus* st,
                       double dtv, double dtx, const VerletState& V, hipStream_t stream) {
  if (sp.n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_move, dim3(nblk(sp.n, 256)), dim3(256), 0, stream, P, S, G, sp, st, dtv, dtx, V);
  return hipGetLastError();
}

__global__ void k_clear_ints(int* p, int n) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0;
}

hipError_t launch_clear(int* p, int n, hipStream_t stream) {
  hipLaunchKernelGGL(k_clear_ints, dim3(nblk(n, 64)), dim3(64), 0, stream, p, n);
  return hipGetLastError();
}

hipError_t launch_key(const Particles& P, const Sorted& S, const Grid& G, const StepParams& sp, DevStatus* st,
                      const VerletState& V, hipStream_t stream) {
  if (sp.n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_key, dim3(nblk(sp.n, 256)), dim3(256), 0, stream, P, S, G, sp, st, V);
  return hipGetLastError();
}

// Blocks of a gated launch (grid-stride loops): enough to fill the chip.
constexpr int kGatedBlocks = 2048;

hipError_t launch_bin(const Particles& P, const Sorted& S, const Grid& G, int n, int cx, hipStream_t stream,
                      const int* gate, int stamp) {
  int nb = nblk(G.ncells, kScanTile);
  hipLaunchKernelGGL(k_scan_reduce, dim3(nb), dim3(kScanThreads), 0, stream, S.count, G.ncells, S.block_sums, gate,
                     stamp);
  if (nb <= kScanFoldTiles) {
    hipLaunchKernelGGL(k_scan_down<true>, dim3(nb), dim3(kScanThreads), 0, stream, S.count, G.ncells, S.block_sums,
                       S.start, n, gate, stamp);
  } else {
    hipLaunchKernelGGL(k_scan_top, dim3(1), dim3(kScanThreads), 0, stream, S.block_sums, nb, gate, stamp);
    hipLaunchKernelGGL(k_scan_down<false>, dim3(nb), dim3(kScanThreads), 0, stream, S.count, G.ncells, S.block_sums,
                       S.start, n, gate, stamp);
  }
  // gated launches mostly find nothing to do: small grids that stride
  const int cap = gate ? kGatedBlocks : (1 << 30);
  if (n > 0) hipLaunchKernelGGL(k_scatter, dim3(std::min(nblk(n, 256), cap)), dim3(256), 0, stream, S, n, gate, stamp);
  hipLaunchKernelGGL(k_canon, dim3(std::min(nblk(G.ncells, 256), cap)), dim3(256), 0, stream, P, S, G.ncells, cx, n,
                     gate, stamp);
  return hipGetLastError();
}

// Verlet list build: cell keys on the list's grid (and the build positions),
// binning, then the candidates of every particle within the contact
// distance plus the skin.  Gated on the stale flag unless stamp == 0.
__global__ __launch_bounds__(256) void k_vkey(Particles P, Sorted S, VerletState V, int n, int cx, DevStatus* st,
                                              int stamp) {
  const int* gate = stamp ? &st->rebuild : nullptr;
  if (gate && *gate != stamp) return;  // Verlet list still valid: nothing to rebuild
  if (blockIdx.x == 0 && threadIdx.x == 0) atomicAdd(&st->n_rebuild, 1);
  const Grid& G = V.Gv;
  for (int s = blockIdx.x * blockDim.x + threadIdx.x; s < n; s += gridDim.x * blockDim.x) {
    const double x[3] = {P.x[cx][0][s], P.x[cx][1][s], P.x[cx][2][s]};
    int c[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      c[a] = min(max(ref_cell(x[a], G, a), 0), G.n[a] - 1);
      V.xb[a][s] = x[a];
    }
    const int key = (c[1] * G.n[2] + c[2]) * G.n[0] + c[0];
    S.key[s] = key;
    S.slot[s] = atomicAdd(&S.count[key], 1);
  }
}

__global__ __launch_bounds__(256) void k_vdetect(Sorted S, VerletState V, int n, NbrList L, const int* gate,
                                                 int stamp, const DevStatus* st) {
  if (gate && *gate != stamp) return;  // Verlet list still valid: nothing to rebuild
  if (st->code) return;                // an error ended the run
  const int bs = blockDim.x;
  for (int blk = xcd_tile(blockIdx.x, gridDim.x); blk * bs < n; blk += gridDim.x) {
    const int k = blk * bs + threadIdx.x;
    if (k >= n) break;
    const int s = S.order[k];
    if ((unsigned)s >= (unsigned)n) continue;
    const double* pr = S.posr + 4 * (size_t)k;
    const d3 x{pr[0], pr[1], pr[2]};
    const double r = pr[3] + V.skin;
    int c = 0;
    for_each_candidate(S, V.Gv, 0, 0, n, x, [&](int k2, int ix, int iy, int iz) {
      if (k2 == k) return;
      const double* q2 = S.posr + 4 * (size_t)k2;
      d3 xo{ix ? q2[0] + image_shift(V.Gv, 0, ix) : q2[0], iy ? q2[1] + image_shift(V.Gv, 1, iy) : q2[1],
            iz ? q2[2] + image_shift(V.Gv, 2, iz) : q2[2]};
      d3 AB = xo - x;
      double s2 = AB.x * AB.x + AB.y * AB.y + AB.z * AB.z;
      double R = r + q2[3];
      if (s2 > R * R * 1.000000000001) return;
      if (c < L.cap) L.ent[(size_t)c * n + s] = (unsigned)S.order[k2];
      ++c;
    });
    L.count[s] = c;
  }
}

hipError_t launch_verlet_build(const Particles& P, const Sorted& S, const VerletState& V, int n, int cx,
                               DevStatus* st, const NbrList& L, int stamp, hipStream_t stream) {
  if (n == 0) return hipSuccess;
  const int nb = stamp ? std::min(nblk(n, 256), kGatedBlocks) : nblk(n, 256);
  hipLaunchKernelGGL(k_vkey, dim3(nb), dim3(256), 0, stream, P, S, V, n, cx, st, stamp);
  const int* gate = stamp ? &st->rebuild : nullptr;
  hipError_t e = launch_bin(P, S, V.Gv, n, cx, stream, gate, stamp);
  if (e != hipSuccess) return e;
  // one wave per workgroup (the candidate loops differ in length from
  // particle to particle)
  const int nb1 = stamp ? std::min(nblk(n, 64), 4 * kGatedBlocks) : nblk(n, 64);
  hipLaunchKernelGGL(k_vdetect, dim3(nb1), dim3(64), 0, stream, S, V, n, L, gate, stamp, st);
  return hipGetLastError();
}

int scan_blocks(int ncells) { return nblk(ncells, kScanTile); }

template <int MODE, bool MIXED>
static void launch_forces_t(const Particles& P, const Sorted& S, const Grid& G, const BoxObstacle* obs,
                            const HookeModel* models, const StepParams& sp, DevStatus* st, int* pairs,
                            const NbrList& L, const VerletState& V, const NbrList& CL, hipStream_t stream) {
  if (P.shape)
    launch_forces_shp<MODE, MIXED>(P, S, G, obs, models, sp, st, pairs, L, V, CL, stream);
  else
    launch_forces_sph<MODE, MIXED>(P, S, G, obs, models, sp, st, pairs, L, V, CL, stream);
}

hipError_t launch_vfilter(const Particles& P, const Sorted& S, const Grid& G, const VerletState& V, int n, int cx,
                          const NbrList& L, const NbrList& CL, DevStatus* st, hipStream_t stream) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_vfilter, dim3(nblk(n, 256)), dim3(256), 0, stream, P, S, G, V, n, cx, L, CL, st);
  return hipGetLastError();
}

hipError_t launch_detect(const Sorted& S, const Grid& G, int n, const NbrList& L, const DevStatus* st,
                         hipStream_t stream, int* mark) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_detect, dim3(nblk(n, 256)), dim3(256), 0, stream, S, G, n, L, mark, st);
  return hipGetLastError();
}

hipError_t launch_forces(int mode, bool mixed, const Particles& P, const Sorted& S, const Grid& G,
                         const BoxObstacle* obs, const HookeModel* models, const StepParams& sp, DevStatus* st,
                         int* pairs, const NbrList& L, const VerletState& V, const NbrList& CL,
                         hipStream_t stream) {
  if (sp.n == 0) return hipSuccess;
  if (mixed) {
    if (mode == kForcesOnly) launch_forces_t<kForcesOnly, true>(P, S, G, obs, models, sp, st, pairs, L, V, CL, stream);
    else if (mode == kForcesKick) launch_forces_t<kForcesKick, true>(P, S, G, obs, models, sp, st, pairs, L, V, CL, stream);
    else launch_forces_t<kFused, true>(P, S, G, obs, models, sp, st, pairs, L, V, CL, stream);
  } else {
    if (mode == kForcesOnly) launch_forces_t<kForcesOnly, false>(P, S, G, obs, models, sp, st, pairs, L, V, CL, stream);
    else if (mode == kForcesKick) launch_forces_t<kForcesKick, false>(P, S, G, obs, models, sp, st, pairs, L, V, CL, stream);
    else launch_forces_t<kFused, false>(P, S, G, obs, models, sp, st, pairs, L, V, CL, stream);
  }
  return hipGetLastError();
}

}  // namespace dem
