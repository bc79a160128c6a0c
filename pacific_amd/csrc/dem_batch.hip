// Batched ensembles: many independent DEM systems (unevenly sized, each with
// its own box, obstacles and contact parameters) advanced together, in one
// of two launch shapes the autotuner chooses between.
//
// Persistent mode (workgroup size 64..512): ONE launch; workgroup b owns
// system sys_order[b] and runs every phase of every step for it, separated by
// workgroup barriers only: no workgroup ever waits on another, so the launch
// cannot deadlock whatever the residency.
//
// The phases reuse the single-system device functions (dem_device.hpp), in
// the same order, so system i of a batch evolves bit for bit like system i
// run on its own.  Per step (SecondOrderLeapFrog, Grains.cpp:344-364):
//   move (half kick + drift + periodic swap + key + atomic slot)
//   scan of the system's cell counts into its start table
//   scatter, per-cell id ordering (counts reset)
//   forces (27-cell sweep + obstacles, gathered per particle)
//   second half kick
// The reference runs one system per process (GrainsExec, ContactBuilderFactory
// and Cell hold static state); the batch is this build's own capability.
//
// Flat mode (workgroup size 0, the default): the same phases as one launch
// each over the particles of every system at once -- each thread reads its
// system's grid, obstacles and models from the system table -- with the cell
// counts of all systems scanned as one array (the systems' cells and
// particles are laid out in the same order, so a global exclusive scan is
// each system's own start table offset by its first particle), and a
// per-step candidate list (k_fb_detect) that the force sweep reads.  Large
// systems then spread over the whole chip instead of one workgroup's CU, at
// the price of a few launches per step; the forces, the second half kick
// and the next drift share one kernel, velocities double buffered (the sweep
// reads the neighbours' while each thread writes its own).  Same device
// functions, same contact order: members evolve bit for bit as in the
// persistent mode.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/grains_gpu.h"
#include "dem_device.hpp"
#include "dem_launch.hpp"
#include "dem_system.hpp"

namespace dem {

// candidates per particle of the flat mode's per-step list (a particle with
// more is searched in its 27 cells by the force sweep)
constexpr int kBatchNbrCap = 32;

struct BatchSys {
  Grid grid;
  double g[3];
  double dt;
  int p0, np;    // particle range (storage index = sorted slot range)
  int c0;        // first global cell key
  int s0;        // first entry of the start table (c0 + system index)
  int o0, nobs;  // obstacle range
  int m0;        // material table base
  int pad;
};

// Exclusive scan of count[c0 .. c0+nc) into start[s0 .. s0+nc], offset by
// p0, by one workgroup; lds holds blockDim ints.
__device__ void wg_scan(const int* __restrict__ count, int* __restrict__ start, int c0, int s0, int nc, int p0,
                        int* lds) {
  const int t = threadIdx.x, nt = blockDim.x;
  int carry = p0;
  for (int base = 0; base < nc; base += nt) {
    int i = base + t;
    int v = i < nc ? count[c0 + i] : 0;
    lds[t] = v;
    __syncthreads();
    for (int off = 1; off < nt; off <<= 1) {
      int add = t >= off ? lds[t - off] : 0;
      __syncthreads();
      lds[t] += add;
      __syncthreads();
    }
    if (i < nc) start[s0 + i] = carry + lds[t] - v;
    int total = lds[nt - 1];
    __syncthreads();
    carry += total;
  }
  if (t == 0) start[s0 + nc] = carry;
}

__device__ __forceinline__ StepParams batch_params(const BatchSys& B, int hi) {
  StepParams sp;
  sp.g[0] = B.g[0];
  sp.g[1] = B.g[1];
  sp.g[2] = B.g[2];
  sp.dt = B.dt;
  sp.n = hi;
  sp.cur = 0;
  sp.nobs = B.nobs;
  sp.nmat = kMaxMaterials;
  sp.record_pairs = 0;
  sp.pair_cap = 0;
  sp.scheme = kLeapFrog;
  sp.first_step = 0;
  sp.stats = nullptr;
  sp.mem = ContactMemory{};
  sp.cf = ContactFeat{};
  sp.fsum = nullptr;
  sp.hyd = nullptr;
  return sp;
}

// mode 0: bin the current positions and compute the initial forces (LeapFrog
// priming, Grains.cpp:169-174); mode 1: nsteps time steps.
template <bool MIXED>
__global__ __launch_bounds__(512) void k_batch(const BatchSys* __restrict__ sys, const int* __restrict__ sys_order,
                                               Particles P, Sorted S, const BoxObstacle* __restrict__ obs,
                                               const HookeModel* __restrict__ models, int mode, int nsteps,
                                               DevStatus* st) {
  extern __shared__ int lds[];
  const BatchSys B = sys[sys_order[blockIdx.x]];
  const int t = threadIdx.x, nt = blockDim.x;
  const int lo = B.p0, hi = B.p0 + B.np;
  const int nc = B.grid.ncells;
  StepParams sp;
  sp.g[0] = B.g[0];
  sp.g[1] = B.g[1];
  sp.g[2] = B.g[2];
  sp.dt = B.dt;
  sp.n = hi;
  sp.cur = 0;
  sp.nobs = B.nobs;
  sp.nmat = kMaxMaterials;
  sp.record_pairs = 0;
  sp.pair_cap = 0;
  sp.scheme = kLeapFrog;
  sp.first_step = 0;
  sp.stats = nullptr;
  sp.mem = ContactMemory{};
  sp.cf = ContactFeat{};
  sp.fsum = nullptr;
  sp.hyd = nullptr;
  const BoxObstacle* bobs = obs + B.o0;
  const HookeModel* bmod = models + B.m0;
  const int nrounds = mode == 0 ? 1 : nsteps;
  for (int step = 0; step < nrounds; ++step) {
    // -- move (or, priming, key only) -------------------------------------
    for (int s = lo + t; s < hi; s += nt) {
      Kin k = load_kin(P, s, 0, 0);
      int pid = P.id[s];
      if (mode == 1) {
        double depl = lf_move(k, ld3(P.F, s), ld3(P.T, s), P.mass[s], P.cfac[s], P.inertia[s], 0.5 * B.dt, B.dt);
        if (depl > P.crust[s]) raise_error(st, kMotionError, pid, 0);
      }
      wrap_and_bin(k, B.grid, S, st, s, pid, B.c0);
      st3(P.x[0], s, k.x);
      st3(P.v[0], s, k.v);
      st3(P.w[0], s, k.w);
      st3(P.wb, s, k.wb);
      P.q[0][s] = k.q.v.x;
      P.q[1][s] = k.q.v.y;
      P.q[2][s] = k.q.v.z;
      P.q[3][s] = k.q.w;
    }
    __syncthreads();
    // -- cell starts --------------------------------------------------------
    wg_scan(S.count, S.start, B.c0, B.s0, nc, lo, lds);
    __syncthreads();
    // -- scatter ------------------------------------------------------------
    for (int s = lo + t; s < hi; s += nt) {
      int dst = S.start[B.s0 + (S.key[s] - B.c0)] + S.slot[s];
      if (dst >= lo && dst < hi) S.tmp_order[dst] = s;
    }
    __syncthreads();
    // -- per-cell ordering ----------------------------------------------------
    for (int c = t; c < nc; c += nt) {
      S.count[B.c0 + c] = 0;
      int b = max(S.start[B.s0 + c], lo), e = min(S.start[B.s0 + c + 1], hi);
      canon_cell(P, S, b, e, 0, lo, hi);
    }
    __syncthreads();
    // -- forces ---------------------------------------------------------------
    for (int k = lo + t; k < hi; k += nt) {
      const int s = S.order[k];
      if (s < lo || s >= hi) continue;
      const double* pr = S.posr + 4 * (size_t)k;
      Self me;
      me.x = d3{pr[0], pr[1], pr[2]};
      me.r = pr[3];
      me.id = S.sid[k];
      me.v = ld3(P.v[0], s);
      me.w = ld3(P.w[0], s);
      me.m = P.mass[s];
      me.c = P.crust[s];
      me.req = P.req[s];
      me.mat = P.mat[s];
      const double wf = P.wfac[s];
      d3 F{B.g[0] * wf, B.g[1] * wf, B.g[2] * wf};
      d3 T{0., 0., 0.};
      sweep_particle<MIXED>(P, S, B.grid, bobs, B.nobs, bmod, sp, st, nullptr, B.s0, lo, hi, 0, k, me, F, T);
      st3(P.F, s, F);
      st3(P.T, s, T);
    }
    __syncthreads();
    if (mode == 0) break;
    // -- second half kick ---------------------------------------------------
    for (int s = lo + t; s < hi; s += nt) {
      Kin k;
      k.x = d3{0., 0., 0.};
      k.v = ld3(P.v[0], s);
      k.w = ld3(P.w[0], s);
      k.wb = ld3(P.wb, s);
      k.q = quat{d3{P.q[0][s], P.q[1][s], P.q[2][s]}, P.q[3][s]};
      lf_kick(k, ld3(P.F, s), ld3(P.T, s), P.mass[s], P.cfac[s], P.inertia[s], 0.5 * B.dt);
      st3(P.v[0], s, k.v);
      st3(P.w[0], s, k.w);
      st3(P.wb, s, k.wb);
    }
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// Flat mode kernels.  sys holds the system table with s0 = c0 (one start
// table for all cells); sysid the system of each storage index.

// move = true: first half kick and drift from the stored torsor (the first
// step of a batch); false: the key of the current positions only (priming).
__global__ __launch_bounds__(256) void k_fb_move(const BatchSys* __restrict__ sys, const int* __restrict__ sysid,
                                                 Particles P, Sorted S, int n, bool move, DevStatus* st) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  const BatchSys& B = sys[sysid[s]];
  Kin k = load_kin(P, s, 0, 0);
  const int pid = P.id[s];
  if (move) {
    const double depl = lf_move(k, ld3(P.F, s), ld3(P.T, s), P.mass[s], P.cfac[s], P.inertia[s], 0.5 * B.dt, B.dt);
    if (depl > P.crust[s]) raise_error(st, kMotionError, pid, 0);
  }
  wrap_and_bin(k, B.grid, S, st, s, pid, B.c0);
  st3(P.x[0], s, k.x);
  st3(P.v[0], s, k.v);
  st3(P.w[0], s, k.w);
  st3(P.wb, s, k.wb);
  P.q[0][s] = k.q.v.x;
  P.q[1][s] = k.q.v.y;
  P.q[2][s] = k.q.v.z;
  P.q[3][s] = k.q.w;
}

// One thread per sorted slot: the candidates of the particle in its
// system's 27 cells that pass the narrow phase's conservative distance
// reject, in the order of the cell sweep (k_detect with the system's grid).
__global__ __launch_bounds__(256) void k_fb_detect(const BatchSys* __restrict__ sys, const int* __restrict__ sysid,
                                                   Sorted S, int n, NbrList L) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const int s = S.order[k];
  if ((unsigned)s >= (unsigned)n) {
    L.count[k] = 0;
    return;
  }
  const BatchSys& B = sys[sysid[s]];
  const double* pr = S.posr + 4 * (size_t)k;
  const d3 x{pr[0], pr[1], pr[2]};
  const double r = pr[3];
  int c = 0;
  for_each_candidate(S, B.grid, B.s0, B.p0, B.p0 + B.np, x, [&](int k2, int ix, int iy, int iz) {
    if (k2 == k) return;
    const double* q2 = S.posr + 4 * (size_t)k2;
    const d3 xo{ix ? q2[0] + image_shift(B.grid, 0, ix) : q2[0], iy ? q2[1] + image_shift(B.grid, 1, iy) : q2[1],
                iz ? q2[2] + image_shift(B.grid, 2, iz) : q2[2]};
    const d3 AB = xo - x;
    const double s2 = AB.x * AB.x + AB.y * AB.y + AB.z * AB.z;
    const double R = r + q2[3];
    if (s2 > R * R * 1.000000000001) return;
    if (c < L.cap) L.ent[(size_t)c * n + k] = nbr_entry(k2, ix, iy, iz);
    ++c;
  });
  L.count[k] = c;
}

// One thread per sorted slot: the particle's contacts (its candidates of L,
// or its system's 27 cells when the candidates overflowed the list) and
// obstacles, then MODE 0 the torsor only (priming), 1 the second
// half kick (last step of a batch; torsor stored), 2 the second half kick,
// the next first half kick and drift and the next cell key.  Velocities are
// read from buffer cv and written to cv ^ 1.
// (held to three waves per SIMD: the FP64 variant would take two)
template <bool MIXED, int MODE>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3))) void k_fb_forces(const BatchSys* __restrict__ sys, const int* __restrict__ sysid,
                                                   Particles P, Sorted S, const BoxObstacle* __restrict__ obs,
                                                   const HookeModel* __restrict__ models, int n, int cv,
                                                   DevStatus* st, NbrList L) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const int s = S.order[k];
  if ((unsigned)s >= (unsigned)n) return;
  const BatchSys& B = sys[sysid[s]];
  const int lo = B.p0, hi = B.p0 + B.np;
  const StepParams sp = batch_params(B, hi);
  const double* pr = S.posr + 4 * (size_t)k;
  Self me;
  me.x = d3{pr[0], pr[1], pr[2]};
  me.r = pr[3];
  me.id = S.sid[k];
  me.v = ld3(P.v[cv], s);
  me.w = ld3(P.w[cv], s);
  me.m = P.mass[s];
  me.c = P.crust[s];
  me.req = P.req[s];
  me.mat = P.mat[s];
  const double wf = P.wfac[s];
  d3 F{B.g[0] * wf, B.g[1] * wf, B.g[2] * wf};
  d3 T{0., 0., 0.};
  sweep_particle<MIXED>(P, S, B.grid, obs + B.o0, B.nobs, models + B.m0, sp, st, nullptr, B.s0, lo, hi, cv, k, me, F,
                        T, nullptr, 0, &L);
  if (MODE != 2) {
    st3(P.F, s, F);
    st3(P.T, s, T);
  }
  if (MODE == 0) return;
  Kin kin;
  kin.x = me.x;
  kin.v = me.v;
  kin.w = me.w;
  kin.wb = ld3(P.wb, s);
  kin.q = quat{d3{P.q[0][s], P.q[1][s], P.q[2][s]}, P.q[3][s]};
  const double cf = P.cfac[s], I = P.inertia[s];
  lf_kick(kin, F, T, me.m, cf, I, 0.5 * B.dt);
  if (MODE == 2) {
    const double depl = lf_move(kin, F, T, me.m, cf, I, 0.5 * B.dt, B.dt);
    if (depl > me.c) raise_error(st, kMotionError, me.id, 0);
    wrap_and_bin(kin, B.grid, S, st, s, me.id, B.c0);
    st3(P.x[0], s, kin.x);
    P.q[0][s] = kin.q.v.x;
    P.q[1][s] = kin.q.v.y;
    P.q[2][s] = kin.q.v.z;
    P.q[3][s] = kin.q.w;
  }
  st3(P.v[cv ^ 1], s, kin.v);
  st3(P.w[cv ^ 1], s, kin.w);
  st3(P.wb, s, kin.wb);
}

__global__ __launch_bounds__(256) void k_fb_copy_v(Particles P, int n) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  st3(P.v[0], s, ld3(P.v[1], s));
  st3(P.w[0], s, ld3(P.w[1], s));
}

}  // namespace dem

using namespace dem;

struct grains_gpu_batch {
  int nsys = 0;
  int ntot = 0;
  int precision = GRAINS_FP64;
  int device = 0;
  int block = 0;  // flat mode; 64..512: the persistent launch's workgroup size
  int graph_steps = 0;
  double time = 0.;
  std::vector<int> np, p0;
  std::vector<double> dt;
  std::vector<std::vector<int>> ids;  // per system: id of each storage slot (1..n_i)
  std::string last_error;
  DevStatus last_status{};
  Particles P{};
  Sorted S{};
  BatchSys* d_sys = nullptr;
  int* d_order = nullptr;
  // flat mode: system table with one start table, system of each particle,
  // sorted arrays over that start table, second velocity buffers
  BatchSys* d_fsys = nullptr;
  int* d_sysid = nullptr;
  Sorted Sf{};
  int ncells = 0;
  int* d_nbr_count = nullptr;  // per-step candidate list of the flat mode
  unsigned* d_nbr = nullptr;
  BoxObstacle* d_obs = nullptr;
  HookeModel* d_models = nullptr;
  DevStatus* d_status = nullptr;
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  std::vector<void*> allocations;
  ~grains_gpu_batch();
  int launch(int mode, int nsteps);
  int launch_flat(int mode, int nsteps);
  int check();
};

namespace {

int bfail(grains_gpu_batch* b, int code, const std::string& m) {
  if (b) b->last_error = m;
  return code;
}

#define BCHECK(b, expr)                                                                              \
  do {                                                                                               \
    hipError_t e_ = (expr);                                                                          \
    if (e_ != hipSuccess)                                                                            \
      return bfail(b, GRAINS_DEVICE_ERROR, std::string(#expr) + ": " + hipGetErrorString(e_));       \
  } while (0)

template <typename T>
hipError_t balloc(grains_gpu_batch* b, T** p, size_t n) {
  hipError_t e = hipMalloc((void**)p, std::max<size_t>(n, 1) * sizeof(T));
  if (e == hipSuccess) b->allocations.push_back((void*)*p);
  return e;
}

template <typename T>
hipError_t up(T* dst, const std::vector<T>& v) {
  if (v.empty()) return hipSuccess;
  return hipMemcpy(dst, v.data(), sizeof(T) * v.size(), hipMemcpyHostToDevice);
}

}  // namespace

grains_gpu_batch::~grains_gpu_batch() {
  (void)hipSetDevice(device);
  if (exec) (void)hipGraphExecDestroy(exec);
  if (graph) (void)hipGraphDestroy(graph);
  for (void* p : allocations) (void)hipFree(p);
  if (ev0) (void)hipEventDestroy(ev0);
  if (ev1) (void)hipEventDestroy(ev1);
  if (stream) (void)hipStreamDestroy(stream);
}

int grains_gpu_batch::launch_flat(int mode, int nsteps) {
  const int n = ntot;
  if (n == 0) return GRAINS_OK;
  const dim3 grid((n + 255) / 256), blk(256);
  Grid all{};
  all.ncells = ncells;
  const bool mixed = precision == GRAINS_MIXED;
  const NbrList L{d_nbr_count, d_nbr, kBatchNbrCap, n};
  auto detect = [&]() {
    hipLaunchKernelGGL(k_fb_detect, grid, blk, 0, stream, d_fsys, d_sysid, Sf, n, L);
  };
  auto forces = [&](int m, int cv) {
    if (mixed) {
      if (m == 0) hipLaunchKernelGGL((k_fb_forces<true, 0>), grid, blk, 0, stream, d_fsys, d_sysid, P, Sf, d_obs, d_models, n, cv, d_status, L);
      else if (m == 1) hipLaunchKernelGGL((k_fb_forces<true, 1>), grid, blk, 0, stream, d_fsys, d_sysid, P, Sf, d_obs, d_models, n, cv, d_status, L);
      else hipLaunchKernelGGL((k_fb_forces<true, 2>), grid, blk, 0, stream, d_fsys, d_sysid, P, Sf, d_obs, d_models, n, cv, d_status, L);
    } else {
      if (m == 0) hipLaunchKernelGGL((k_fb_forces<false, 0>), grid, blk, 0, stream, d_fsys, d_sysid, P, Sf, d_obs, d_models, n, cv, d_status, L);
      else if (m == 1) hipLaunchKernelGGL((k_fb_forces<false, 1>), grid, blk, 0, stream, d_fsys, d_sysid, P, Sf, d_obs, d_models, n, cv, d_status, L);
      else hipLaunchKernelGGL((k_fb_forces<false, 2>), grid, blk, 0, stream, d_fsys, d_sysid, P, Sf, d_obs, d_models, n, cv, d_status, L);
    }
  };
  hipLaunchKernelGGL(k_fb_move, grid, blk, 0, stream, d_fsys, d_sysid, P, Sf, n, mode != 0, d_status);
  BCHECK(this, launch_bin(P, Sf, all, n, 0, stream));
  detect();
  if (mode == 0) {
    forces(0, 0);
    BCHECK(this, hipGetLastError());
    return GRAINS_OK;
  }
  int cv = 0;
  for (int j = 1; j < nsteps; ++j) {
    forces(2, cv);
    cv ^= 1;
    BCHECK(this, launch_bin(P, Sf, all, n, 0, stream));
    detect();
  }
  forces(1, cv);
  cv ^= 1;
  if (cv) hipLaunchKernelGGL(k_fb_copy_v, grid, blk, 0, stream, P, n);
  BCHECK(this, hipGetLastError());
  return GRAINS_OK;
}

int grains_gpu_batch::launch(int mode, int nsteps) {
  if (block == 0) return launch_flat(mode, nsteps);
  size_t lds = sizeof(int) * (size_t)block;
  if (precision == GRAINS_MIXED)
    hipLaunchKernelGGL(k_batch<true>, dim3(nsys), dim3(block), lds, stream, d_sys, d_order, P, S, d_obs, d_models,
                       mode, nsteps, d_status);
  else
    hipLaunchKernelGGL(k_batch<false>, dim3(nsys), dim3(block), lds, stream, d_sys, d_order, P, S, d_obs, d_models,
                       mode, nsteps, d_status);
  BCHECK(this, hipGetLastError());
  return GRAINS_OK;
}

int grains_gpu_batch::check() {
  DevStatus st;
  BCHECK(this, hipMemcpyAsync(&st, d_status, sizeof(st), hipMemcpyDeviceToHost, stream));
  BCHECK(this, hipStreamSynchronize(stream));
  last_status = st;
  if (st.code) {
    static const char* names[] = {"", "ContactError", "MotionError", "SimulationError"};
    return bfail(this, st.code, std::string(names[st.code & 3]) + " raised on the device (bodies " +
                                    std::to_string(st.id0) + " and " + std::to_string(st.id1) + ")");
  }
  return GRAINS_OK;
}

extern "C" {

int grains_gpu_batch_create(grains_gpu_system* const* systems, int nsys, int precision, grains_gpu_batch** out) {
  if (!systems || nsys <= 0 || !out) return GRAINS_ARGUMENT_ERROR;
  grains_gpu_batch* b = new grains_gpu_batch();
  *out = b;
  b->nsys = nsys;
  b->precision = precision;
  b->device = systems[0]->cfg.device;
  std::vector<BatchSys> hs((size_t)nsys);
  std::vector<BoxObstacle> hobs;
  std::vector<HookeModel> hmod;
  std::vector<double> hx[3], hv[3], hw[3], hwb[3], hq[4], hrad, hmass, hI, hcr, hwf, hcf, hreq;
  std::vector<int> hid, hmat;
  long cells = 0;
  for (int i = 0; i < nsys; ++i) {
    grains_gpu_system* s = systems[i];
    if (s->initialized) return bfail(b, GRAINS_ARGUMENT_ERROR, "batch members must not be initialised");
    if (s->has_memory()) return bfail(b, GRAINS_CONFIG_ERROR, "batch members use contact models without memory");
    if (s->cfg.scheme != GRAINS_LEAPFROG) return bfail(b, GRAINS_CONFIG_ERROR, "batches run SecondOrderLeapFrog");
    if (s->has_shapes()) return bfail(b, GRAINS_CONFIG_ERROR, "batch members hold spheres only");
    // model checks and grid of the member (same rules as one system)
    std::vector<int> used(kMaxMaterials, 0);
    for (int m : s->h_mat) used[m] |= 1;
    for (auto& o : s->obstacles) used[o.mat] |= 2;
    for (int a = 0; a < kMaxMaterials; ++a)
      for (int c = 0; c < kMaxMaterials; ++c)
        if ((used[a] & 1) && used[c] && !s->models[a * kMaxMaterials + c].defined) {
          if (s->has_default) s->models[a * kMaxMaterials + c] = s->default_model;
          else return bfail(b, GRAINS_CONFIG_ERROR, "member " + std::to_string(i) + " lacks a contact model");
        }
    int e = s->build_grid();
    if (e) return bfail(b, e, "member " + std::to_string(i) + ": " + s->last_error);
    BatchSys& B = hs[i];
    B.grid = s->grid;
    for (int a = 0; a < 3; ++a) B.g[a] = s->cfg.gravity[a];
    B.dt = s->cfg.dt;
    B.p0 = b->ntot;
    B.np = s->nparticles;
    B.c0 = (int)cells;
    B.s0 = (int)cells + i;
    B.o0 = (int)hobs.size();
    B.nobs = (int)s->obstacles.size();
    B.m0 = (int)hmod.size();
    B.pad = 0;
    cells += s->grid.ncells;
    if (cells > (1L << 30)) return bfail(b, GRAINS_CONFIG_ERROR, "too many cells in the batch");
    for (auto& o : s->obstacles) hobs.push_back(o);
    for (int m = 0; m < kMaxMaterials * kMaxMaterials; ++m) hmod.push_back(s->models[m]);
    for (int p = 0; p < s->nparticles; ++p) {
      for (int a = 0; a < 3; ++a) {
        hx[a].push_back(s->h_x[3 * p + a]);
        hv[a].push_back(s->h_v[3 * p + a]);
        hw[a].push_back(s->h_w[3 * p + a]);
        hwb[a].push_back(s->h_wb[3 * p + a]);
      }
      for (int a = 0; a < 4; ++a) hq[a].push_back(s->h_q[4 * p + a]);
      hrad.push_back(s->h_rad[p]);
      hmass.push_back(s->h_mass[p]);
      hI.push_back(s->h_inertia[p]);
      hcr.push_back(s->h_crust[p]);
      hwf.push_back(s->h_wfac[p]);
      hcf.push_back(s->h_cfac[p]);
      hreq.push_back(s->h_req[p]);
      hid.push_back(p + 1);
      hmat.push_back(s->h_mat[p]);
    }
    b->np.push_back(s->nparticles);
    b->p0.push_back(B.p0);
    b->dt.push_back(s->cfg.dt);
    b->ntot += s->nparticles;
  }
  // workgroups take systems largest first
  std::vector<int> order((size_t)nsys);
  for (int i = 0; i < nsys; ++i) order[i] = i;
  std::stable_sort(order.begin(), order.end(), [&](int a, int c) { return b->np[a] > b->np[c]; });

  const size_t n = (size_t)b->ntot;
  BCHECK(b, hipSetDevice(b->device));
  BCHECK(b, hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking));
  BCHECK(b, hipEventCreate(&b->ev0));
  BCHECK(b, hipEventCreate(&b->ev1));
  Particles& P = b->P;
  for (int a = 0; a < 3; ++a) {
    BCHECK(b, balloc(b, &P.x[0][a], n));
    BCHECK(b, balloc(b, &P.v[0][a], n));
    BCHECK(b, balloc(b, &P.w[0][a], n));
    P.x[1][a] = P.x[0][a];
    P.v[1][a] = P.v[0][a];
    P.w[1][a] = P.w[0][a];
    BCHECK(b, balloc(b, &P.wb[a], n));
    BCHECK(b, balloc(b, &P.F[a], n));
    BCHECK(b, balloc(b, &P.T[a], n));
    BCHECK(b, up(P.x[0][a], hx[a]));
    BCHECK(b, up(P.v[0][a], hv[a]));
    BCHECK(b, up(P.w[0][a], hw[a]));
    BCHECK(b, up(P.wb[a], hwb[a]));
  }
  for (int a = 0; a < 4; ++a) {
    BCHECK(b, balloc(b, &P.q[a], n));
    BCHECK(b, up(P.q[a], hq[a]));
  }
  BCHECK(b, balloc(b, &P.rad, n));
  BCHECK(b, balloc(b, &P.mass, n));
  BCHECK(b, balloc(b, &P.inertia, n));
  BCHECK(b, balloc(b, &P.crust, n));
  BCHECK(b, balloc(b, &P.wfac, n));
  BCHECK(b, balloc(b, &P.cfac, n));
  BCHECK(b, balloc(b, &P.req, n));
  BCHECK(b, balloc(b, &P.id, n));
  BCHECK(b, balloc(b, &P.mat, n));
  BCHECK(b, up(P.rad, hrad));
  BCHECK(b, up(P.mass, hmass));
  BCHECK(b, up(P.inertia, hI));
  BCHECK(b, up(P.crust, hcr));
  BCHECK(b, up(P.wfac, hwf));
  BCHECK(b, up(P.cfac, hcf));
  BCHECK(b, up(P.req, hreq));
  BCHECK(b, up(P.id, hid));
  BCHECK(b, up(P.mat, hmat));
  Sorted& S = b->S;
  BCHECK(b, balloc(b, &S.key, n));
  BCHECK(b, balloc(b, &S.slot, n));
  BCHECK(b, balloc(b, &S.count, (size_t)cells));
  BCHECK(b, balloc(b, &S.start, (size_t)cells + nsys));
  BCHECK(b, balloc(b, &S.tmp_order, n));
  BCHECK(b, balloc(b, &S.order, n));
  BCHECK(b, balloc(b, &S.sid, n));
  BCHECK(b, balloc(b, &S.posr, 4 * n));
  S.block_sums = nullptr;
  BCHECK(b, hipMemset(S.count, 0, sizeof(int) * (size_t)std::max<long>(cells, 1)));
  BCHECK(b, hipMemset(S.start, 0, sizeof(int) * ((size_t)cells + nsys)));
  BCHECK(b, hipMemset(S.tmp_order, 0, sizeof(int) * n));
  BCHECK(b, hipMemset(S.order, 0, sizeof(int) * n));
  BCHECK(b, balloc(b, &b->d_sys, (size_t)nsys));
  BCHECK(b, up(b->d_sys, hs));
  // flat mode
  {
    std::vector<BatchSys> fs = hs;
    for (auto& B : fs) B.s0 = B.c0;
    BCHECK(b, balloc(b, &b->d_fsys, (size_t)nsys));
    BCHECK(b, up(b->d_fsys, fs));
    std::vector<int> sid(n);
    for (int i = 0; i < nsys; ++i)
      for (int p = 0; p < b->np[i]; ++p) sid[(size_t)b->p0[i] + p] = i;
    BCHECK(b, balloc(b, &b->d_sysid, n));
    BCHECK(b, up(b->d_sysid, sid));
    b->ncells = (int)cells;
    b->Sf = S;
    BCHECK(b, balloc(b, &b->Sf.start, (size_t)cells + 1));
    BCHECK(b, hipMemset(b->Sf.start, 0, sizeof(int) * ((size_t)cells + 1)));
    BCHECK(b, balloc(b, &b->Sf.block_sums, (size_t)scan_blocks((int)cells) + 1));
    for (int a = 0; a < 3; ++a) {
      BCHECK(b, balloc(b, &P.v[1][a], n));
      BCHECK(b, balloc(b, &P.w[1][a], n));
    }
    BCHECK(b, balloc(b, &b->d_nbr_count, n));
    BCHECK(b, balloc(b, &b->d_nbr, (size_t)kBatchNbrCap * n));
  }
  BCHECK(b, balloc(b, &b->d_order, (size_t)nsys));
  BCHECK(b, up(b->d_order, order));
  BCHECK(b, balloc(b, &b->d_obs, hobs.size()));
  BCHECK(b, up(b->d_obs, hobs));
  BCHECK(b, balloc(b, &b->d_models, hmod.size()));
  BCHECK(b, up(b->d_models, hmod));
  BCHECK(b, balloc(b, &b->d_status, 1));
  BCHECK(b, hipMemset(b->d_status, 0, sizeof(DevStatus)));
  b->time = systems[0]->cfg.t_start;
  int e = b->launch(0, 0);  // bin + initial forces
  if (e) return e;
  return b->check();
}

void grains_gpu_batch_destroy(grains_gpu_batch* b) { delete b; }

int grains_gpu_batch_set_block(grains_gpu_batch* b, int block) {
  if (!b || (block != 0 && (block < 64 || block > 512 || block % 64))) return GRAINS_ARGUMENT_ERROR;
  if (block != b->block && b->exec) {
    (void)hipGraphExecDestroy(b->exec);
    (void)hipGraphDestroy(b->graph);
    b->exec = nullptr;
    b->graph = nullptr;
    b->graph_steps = 0;
  }
  b->block = block;
  return GRAINS_OK;
}

// nsteps time steps of every member.  Whole chunks of `chunk` steps replay a
// captured hipGraph of one persistent launch; chunk <= 0 launches directly.
int grains_gpu_batch_step(grains_gpu_batch* b, int nsteps, int chunk) {
  if (!b || nsteps < 0) return GRAINS_ARGUMENT_ERROR;
  BCHECK(b, hipSetDevice(b->device));
  int done = 0;
  if (chunk > 0 && nsteps >= chunk) {
    if (!b->exec || b->graph_steps != chunk) {
      if (b->exec) {
        (void)hipGraphExecDestroy(b->exec);
        (void)hipGraphDestroy(b->graph);
        b->exec = nullptr;
      }
      BCHECK(b, hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
      int e = b->launch(1, chunk);
      hipGraph_t g = nullptr;
      hipError_t ec = hipStreamEndCapture(b->stream, &g);
      if (e) return e;
      BCHECK(b, ec);
      b->graph = g;
      BCHECK(b, hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
      b->graph_steps = chunk;
    }
    for (; done + chunk <= nsteps; done += chunk) BCHECK(b, hipGraphLaunch(b->exec, b->stream));
  }
  if (done < nsteps) {
    int e = b->launch(1, nsteps - done);
    if (e) return e;
  }
  for (int i = 0; i < nsteps; ++i) b->time += b->dt[0];
  return b->check();
}

int grains_gpu_batch_step_timed(grains_gpu_batch* b, int nsteps, int chunk, double* ms) {
  if (!b || !ms) return GRAINS_ARGUMENT_ERROR;
  BCHECK(b, hipSetDevice(b->device));
  // build the graph outside the timed region
  if (chunk > 0 && nsteps >= chunk && (!b->exec || b->graph_steps != chunk)) {
    int e = grains_gpu_batch_step(b, 0, chunk);
    if (e) return e;
    BCHECK(b, hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
    e = b->launch(1, chunk);
    hipGraph_t g = nullptr;
    hipError_t ec = hipStreamEndCapture(b->stream, &g);
    if (e) return e;
    BCHECK(b, ec);
    if (b->exec) {
      (void)hipGraphExecDestroy(b->exec);
      (void)hipGraphDestroy(b->graph);
    }
    b->graph = g;
    BCHECK(b, hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
    b->graph_steps = chunk;
  }
  BCHECK(b, hipEventRecord(b->ev0, b->stream));
  int done = 0;
  if (chunk > 0)
    for (; done + chunk <= nsteps; done += chunk) BCHECK(b, hipGraphLaunch(b->exec, b->stream));
  if (done < nsteps) {
    int e = b->launch(1, nsteps - done);
    if (e) return e;
  }
  BCHECK(b, hipEventRecord(b->ev1, b->stream));
  BCHECK(b, hipEventSynchronize(b->ev1));
  float t = 0.f;
  BCHECK(b, hipEventElapsedTime(&t, b->ev0, b->ev1));
  *ms = t;
  for (int i = 0; i < nsteps; ++i) b->time += b->dt[0];
  return b->check();
}

// Times `trial_steps` steps in every launch shape and keeps the fastest;
// ms_out[i] receives the time of candidate i: 0 (flat mode), then the
// persistent launch with workgroups of 64, 128, ..., 512 lanes (whole waves,
// so that sizes between the powers of two can fill the wave slots a CU has
// at the kernel's register count).
int grains_gpu_batch_autotune(grains_gpu_batch* b, int trial_steps, int* best_block, double* ms_out) {
  if (!b || trial_steps <= 0) return GRAINS_ARGUMENT_ERROR;
  BCHECK(b, hipSetDevice(b->device));
  constexpr int kCand = 9;
  double best = 1e300;
  int bestb = b->block;
  for (int i = 0; i < kCand; ++i) {
    const int cand[kCand] = {0, 64, 128, 192, 256, 320, 384, 448, 512};
    grains_gpu_batch_set_block(b, cand[i]);
    BCHECK(b, hipEventRecord(b->ev0, b->stream));
    int e = b->launch(1, trial_steps);
    if (e) return e;
    BCHECK(b, hipEventRecord(b->ev1, b->stream));
    BCHECK(b, hipEventSynchronize(b->ev1));
    float t = 0.f;
    BCHECK(b, hipEventElapsedTime(&t, b->ev0, b->ev1));
    for (int k = 0; k < trial_steps; ++k) b->time += b->dt[0];
    if (ms_out) ms_out[i] = t;
    if (t < best) {
      best = t;
      bestb = cand[i];
    }
  }
  grains_gpu_batch_set_block(b, bestb);
  if (best_block) *best_block = bestb;
  return b->check();
}

size_t grains_gpu_batch_num_particles(const grains_gpu_batch* b, int isys) {
  if (!b) return 0;
  if (isys < 0) return (size_t)b->ntot;
  return isys < b->nsys ? (size_t)b->np[isys] : 0;
}

double grains_gpu_batch_time(const grains_gpu_batch* b) { return b ? b->time : 0.; }

int grains_gpu_batch_get_state(grains_gpu_batch* b, int isys, double* x, double* v, double* w, double* q, double* F,
                               double* T) {
  if (!b || isys < 0 || isys >= b->nsys) return GRAINS_ARGUMENT_ERROR;
  BCHECK(b, hipSetDevice(b->device));
  BCHECK(b, hipStreamSynchronize(b->stream));
  const int n = b->np[isys], p0 = b->p0[isys];
  std::vector<int> ids((size_t)n);
  BCHECK(b, hipMemcpy(ids.data(), b->P.id + p0, sizeof(int) * n, hipMemcpyDeviceToHost));
  std::vector<double> buf((size_t)n);
  auto down = [&](double* dst, double* src, int stride, int comp) -> hipError_t {
    if (!dst) return hipSuccess;
    hipError_t e = hipMemcpy(buf.data(), src + p0, sizeof(double) * n, hipMemcpyDeviceToHost);
    for (int i = 0; i < n; ++i) dst[(size_t)stride * (ids[i] - 1) + comp] = buf[i];
    return e;
  };
  for (int a = 0; a < 3; ++a) {
    BCHECK(b, down(x, b->P.x[0][a], 3, a));
    BCHECK(b, down(v, b->P.v[0][a], 3, a));
    BCHECK(b, down(w, b->P.w[0][a], 3, a));
    BCHECK(b, down(F, b->P.F[a], 3, a));
    BCHECK(b, down(T, b->P.T[a], 3, a));
  }
  for (int a = 0; a < 4; ++a) BCHECK(b, down(q, b->P.q[a], 4, a));
  return b->check();
}

const char* grains_gpu_batch_last_error(const grains_gpu_batch* b) { return b ? b->last_error.c_str() : "null batch"; }

}  // extern "C"
