// Discrete adjoint of the SecondOrderLeapFrog DEM step (spheres, Hooke model):
// gradients of a scalar function of the final state with respect to the
// initial positions, velocities, angular velocities and to (kn, en, mu) of one
// material pair.
//
// Step map (h = dt/2; sphere rotation is isotropic, so the quaternion cancels
// out of the angular update in exact arithmetic):
//   vh = v_n + F_n h/m,  wh = w_n + T_n h/I,  x_{n+1} = x_n + vh dt,
//   (F,T)_{n+1} = f(x_{n+1}, vh, wh),  v_{n+1} = vh + F_{n+1} h/m, ...
// Reverse step n, per particle, with the contact VJP g of f at step n+1:
//   xb' = xb + g_x;  vhb = vb + g_v + xb' dt;  whb = wb + g_w
//   Fb_n = c vhb h/m,  Tb_n = c whb h/I   (c = 2 for n >= 1: F_n also entered
//   the kick that ended step n-1; c = 1 for n = 0)
// and finally the VJP of the initial force evaluation f(x_0, v_0, w_0).
//
// The forward run keeps a full state every K steps; the reverse sweep rebuilds
// each segment from its checkpoint with the same kernels (bit-identical
// states), recording (x_{n+1}, vh_n, wh_n) after every first half kick + drift.
#include <hip/hip_runtime.h>

#include <cmath>
#include <string>
#include <vector>

#include "../../include/grains_gpu.h"
#include "dem_adjoint.hpp"
#include "dem_device.hpp"
#include "dem_launch.hpp"
#include "dem_system.hpp"

namespace dem {

struct AdjArrays {
  double* xb[3];
  double* vb[3];
  double* wb[3];
  double* Fb_in[3];
  double* Tb_in[3];
  double* Fb_out[3];
  double* Tb_out[3];
  double* gth;  // 3 per particle: kn, beta, mu
};

struct AdjVisitor {
  const Particles& P;
  const Sorted& S;
  const BoxObstacle* obs;
  const HookeModel* models;
  const AdjArrays& A;
  int cv;
  const Self& me;
  int ta, tb;  // target material pair
  d3 F0b, T0b;
  PairGrad acc;

  __device__ __forceinline__ bool target(int ma, int mb) const {
    return (ma == ta && mb == tb) || (ma == tb && mb == ta);
  }
  __device__ __forceinline__ ContactIn base(const HookeModel& hm) const {
    ContactIn in;
    in.r0 = me.r;
    in.v0 = me.v;
    in.w0 = me.w;
    in.m0 = me.m;
    in.kn = hm.kn;
    in.beta = hm.beta;
    in.etat = hm.etat;
    in.mu = hm.muc;
    return in;
  }
  __device__ __forceinline__ void pair(int k2, int s2i, int pid2, d3 xo, double ro, d3 Pc, d3 ov, double d) {
    (void)k2;
    (void)Pc;
    (void)ov;
    (void)d;
    const int mo = P.mat[s2i];
    const HookeModel& hm = models[me.mat * kMaxMaterials + mo];
    ContactIn in = base(hm);
    in.r1 = ro;
    in.v1 = ld3(P.v[cv], s2i);
    in.w1 = ld3(P.w[cv], s2i);
    in.m1 = P.mass[s2i];
    PairGrad g;
    pair_vjp(xo - me.x, in, F0b, T0b, ld3(A.Fb_in, s2i), ld3(A.Tb_in, s2i), g);
    acc.x0 += g.x0;
    acc.v0 += g.v0;
    acc.w0 += g.w0;
    // each pair is visited from both sides; its parameter derivative is
    // counted once
    if (target(me.mat, mo) && me.id < pid2) {
      acc.kn += g.kn;
      acc.beta += g.beta;
      acc.mu += g.mu;
    }
  }
  __device__ __forceinline__ void wall(int o, d3 Pc, d3 ov, double d) {
    (void)Pc;
    (void)ov;
    (void)d;
    const BoxObstacle& B = obs[o];
    const HookeModel& hm = models[me.mat * kMaxMaterials + B.mat];
    ContactIn in = base(hm);
    in.r1 = 0.;
    in.v1 = d3{0., 0., 0.};
    in.w1 = d3{0., 0., 0.};
    in.m1 = kObstacleMass;
    PairGrad g;
    wall_vjp(B.B, B.Binv, B.invorigin, B.extent, me.x, in, F0b, T0b, g);
    acc.x0 += g.x0;
    acc.v0 += g.v0;
    acc.w0 += g.w0;
    if (target(me.mat, B.mat)) {
      acc.kn += g.kn;
      acc.beta += g.beta;
      acc.mu += g.mu;
    }
  }
};

// Cotangents of the final state: Fb = vb h/(m c), Tb = wb h/(I c)
__global__ __launch_bounds__(256) void k_adj_init(Particles P, AdjArrays A, int n, double h) {
  int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  const double mc = P.mass[s] * P.cfac[s], Ic = P.inertia[s] * P.cfac[s];
  for (int a = 0; a < 3; ++a) {
    A.Fb_in[a][s] = A.vb[a][s] * h / mc;
    A.Tb_in[a][s] = A.wb[a][s] * h / Ic;
  }
  A.gth[3 * s] = A.gth[3 * s + 1] = A.gth[3 * s + 2] = 0.;
}

// mode 0: reverse step; mode 1: derivative of the initial force evaluation
__global__ __launch_bounds__(256) void k_adj_sweep(Particles P, Sorted S, Grid G, const BoxObstacle* __restrict__ obs,
                                                   int nobs, const HookeModel* __restrict__ models, AdjArrays A, int n,
                                                   int cv, int ta, int tb, double h, double dt, double fac, int mode,
                                                   DevStatus* st) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const int s = S.order[k];
  if ((unsigned)s >= (unsigned)n) return;
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
  AdjVisitor vis{P, S, obs, models, A, cv, me, ta, tb, ld3(A.Fb_in, s), ld3(A.Tb_in, s), PairGrad{}};
  vis.acc.x0 = vis.acc.v0 = vis.acc.w0 = d3{0., 0., 0.};
  vis.acc.kn = vis.acc.beta = vis.acc.mu = 0.;
  for_each_contact(P, S, G, obs, nobs, st, 0, 0, n, k, me, vis);
  A.gth[3 * s] += vis.acc.kn;
  A.gth[3 * s + 1] += vis.acc.beta;
  A.gth[3 * s + 2] += vis.acc.mu;
  d3 xb = ld3(A.xb, s) + vis.acc.x0;
  d3 vb = ld3(A.vb, s) + vis.acc.v0;
  d3 wb = ld3(A.wb, s) + vis.acc.w0;
  if (mode == 0) {
    vb += xb * dt;
    const double mc = P.mass[s] * P.cfac[s], Ic = P.inertia[s] * P.cfac[s];
    st3(A.Fb_out, s, (fac * h / mc) * vb);
    st3(A.Tb_out, s, (fac * h / Ic) * wb);
  }
  st3(A.xb, s, xb);
  st3(A.vb, s, vb);
  st3(A.wb, s, wb);
}

static inline int nblk(int n, int b) { return (n + b - 1) / b; }

// logical state arrays of one checkpoint: x, v, w, wb, F, T (3 each), q (4)
constexpr int kCkptComps = 22;

struct CopyList {
  const double* src[kCkptComps];
  double* dst[kCkptComps];
};

// blockIdx.y selects the array
__global__ __launch_bounds__(256) void k_copy_list(CopyList L, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int a = blockIdx.y;
  if (i < n) L.dst[a][i] = L.src[a][i];
}

// One recorded step: the binned state right after the first half kick and
// the drift -- sorted positions/radii, sorted-slot -> storage index and id,
// the cell-start table, and the half-step velocities in storage order:
//   doubles: posr[4n] | v[3][n] | w[3][n]       ints: order[n] | sid[n] | start[nc1]
struct RecView {
  double* d;
  int* i;
};
inline size_t rec_doubles(size_t n) { return 10 * n; }
inline size_t rec_ints(size_t n, size_t nc1) { return 2 * n + nc1; }

__global__ __launch_bounds__(256) void k_record(Particles P, Sorted S, int cv, int n, int nc1, double* __restrict__ dd,
                                                int* __restrict__ di) {
  const int stride = gridDim.x * blockDim.x;
  const int m = n > nc1 ? n : nc1;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < m; i += stride) {
    if (i < n) {
      const double2* src = reinterpret_cast<const double2*>(S.posr) + 2 * (size_t)i;
      double2* dst = reinterpret_cast<double2*>(dd) + 2 * (size_t)i;
      dst[0] = src[0];
      dst[1] = src[1];
      for (int c = 0; c < 3; ++c) {
        dd[(size_t)(4 + c) * n + i] = P.v[cv][c][i];
        dd[(size_t)(7 + c) * n + i] = P.w[cv][c][i];
      }
      di[i] = S.order[i];
      di[(size_t)n + i] = S.sid[i];
    }
    if (i < nc1) di[2 * (size_t)n + i] = S.start[i];
  }
}

struct AdjointData {
  int nsteps = 0, K = 0, nseg = 0;
  double t0 = 0.;
  long steps0 = 0;
  std::vector<double*> ckpt;  // nseg + 1 (the last one is the end state)
  double* rec_d = nullptr;    // K recorded steps
  int* rec_i = nullptr;
  AdjArrays A{};
  std::vector<void*> allocs;
  ~AdjointData() {
    for (void* p : allocs) (void)hipFree(p);
  }
};

}  // namespace dem

using namespace dem;

namespace {

int afail(grains_gpu_system* s, int code, const std::string& m) {
  s->last_error = m;
  return code;
}

#define ACHECK(s, expr)                                                                             \
  do {                                                                                              \
    hipError_t e_ = (expr);                                                                         \
    if (e_ != hipSuccess)                                                                           \
      return afail(s, GRAINS_DEVICE_ERROR, std::string(#expr) + ": " + hipGetErrorString(e_));      \
  } while (0)

std::vector<double*> state_arrays(grains_gpu_system* s) {
  const int cx = s->cur & 1, cv = (s->cur >> 1) & 1;
  std::vector<double*> a;
  for (int c = 0; c < 3; ++c) a.push_back(s->P.x[cx][c]);
  for (int c = 0; c < 3; ++c) a.push_back(s->P.v[cv][c]);
  for (int c = 0; c < 3; ++c) a.push_back(s->P.w[cv][c]);
  for (int c = 0; c < 3; ++c) a.push_back(s->P.wb[c]);
  for (int c = 0; c < 3; ++c) a.push_back(s->P.F[c]);
  for (int c = 0; c < 3; ++c) a.push_back(s->P.T[c]);
  for (int c = 0; c < 4; ++c) a.push_back(s->P.q[c]);
  return a;
}

int copy_state(grains_gpu_system* s, double* ckpt, bool save) {
  const size_t n = (size_t)s->nparticles;
  auto a = state_arrays(s);
  CopyList L;
  for (int c = 0; c < kCkptComps; ++c) {
    L.src[c] = save ? a[c] : ckpt + c * n;
    L.dst[c] = save ? ckpt + c * n : a[c];
  }
  hipLaunchKernelGGL(k_copy_list, dim3(nblk((int)n, 256), kCkptComps), dim3(256), 0, s->stream, L, (int)n);
  ACHECK(s, hipGetLastError());
  return GRAINS_OK;
}

int save_state(grains_gpu_system* s, double* dst) { return copy_state(s, dst, true); }

int load_state(grains_gpu_system* s, double* src) {
  s->cur = 0;
  int e = copy_state(s, src, false);
  if (e) return e;
  return s->bin_current();
}

}  // namespace

// Recorder hook of grains_gpu_system::enqueue_steps_impl, called right after
// the binning that follows each first half kick + drift
int grains_gpu_system::record_state() {
  if (!recording || !adj) return GRAINS_OK;
  const size_t n = (size_t)nparticles, nc1 = (size_t)grid.ncells + 1;
  if (rec_slot >= adj->K) return afail(this, GRAINS_ARGUMENT_ERROR, "adjoint record buffer exhausted");
  double* dd = adj->rec_d + (size_t)rec_slot * rec_doubles(n);
  int* di = adj->rec_i + (size_t)rec_slot * rec_ints(n, nc1);
  const int m = (int)std::max(n, nc1);
  hipLaunchKernelGGL(k_record, dim3(std::min(nblk(m, 256), 4096)), dim3(256), 0, stream, P, S, (cur >> 1) & 1,
                     (int)n, (int)nc1, dd, di);
  ACHECK(this, hipGetLastError());
  ++rec_slot;
  return GRAINS_OK;
}

extern "C" {

int grains_gpu_adjoint_forward(grains_gpu_system* s, int nsteps, int checkpoint_every) {
  if (!s || !s->initialized || nsteps <= 0 || checkpoint_every <= 0) return GRAINS_ARGUMENT_ERROR;
  if (s->cfg.scheme != GRAINS_LEAPFROG) return afail(s, GRAINS_CONFIG_ERROR, "the adjoint differentiates SecondOrderLeapFrog");
  if (s->cfg.precision != GRAINS_FP64) return afail(s, GRAINS_CONFIG_ERROR, "the adjoint differentiates the FP64 step");
  if (s->has_shapes()) return afail(s, GRAINS_CONFIG_ERROR, "the adjoint differentiates systems of spheres");
  if (s->has_memory())
    return afail(s, GRAINS_CONFIG_ERROR, "the adjoint differentiates the Hooke model without contact memory");
  if (s->verlet)
    return afail(s, GRAINS_CONFIG_ERROR,
                 "the adjoint replays the per-step neighbour search: set the skin to 0 with "
                 "grains_gpu_set_neighbor_list before initialisation");
  for (int m = 0; m < kMaxMaterials * kMaxMaterials; ++m)
    if (s->models[m].defined && s->models[m].kr != 0.)
      return afail(s, GRAINS_CONFIG_ERROR, "the adjoint does not differentiate rolling resistance (kr != 0)");
  ACHECK(s, hipSetDevice(s->device));
  const size_t n = (size_t)s->nparticles;
  const int nseg = (nsteps + checkpoint_every - 1) / checkpoint_every;
  // buffers of a previous run of the same shape are reused
  if (s->adj && (s->adj->nseg != nseg || s->adj->K != checkpoint_every)) s->free_adjoint();
  if (!s->adj) {
    AdjointData* fresh = new AdjointData();
    s->adj = fresh;
    fresh->K = checkpoint_every;
    fresh->nseg = nseg;
    for (int i = 0; i <= nseg; ++i) {
      double* p = nullptr;
      ACHECK(s, hipMalloc((void**)&p, kCkptComps * n * sizeof(double)));
      fresh->allocs.push_back(p);
      fresh->ckpt.push_back(p);
    }
  }
  AdjointData* ad = s->adj;
  ad->nsteps = nsteps;
  ad->t0 = s->time;
  ad->steps0 = s->steps_done;
  for (int seg = 0; seg < ad->nseg; ++seg) {
    int e = save_state(s, ad->ckpt[seg]);
    if (e) return e;
    int len = std::min(checkpoint_every, nsteps - seg * checkpoint_every);
    if ((e = s->enqueue_steps(len))) return e;
  }
  int e = save_state(s, ad->ckpt[ad->nseg]);
  if (e) return e;
  return s->check_status();
}

int grains_gpu_adjoint_backward(grains_gpu_system* s, int matA, int matB, const double* dJ_dx, const double* dJ_dv,
                                const double* dJ_dw, double* grad_x0, double* grad_v0, double* grad_w0,
                                double* grad_params) {
  if (!s || !s->initialized || !s->adj) return GRAINS_ARGUMENT_ERROR;
  if (matA < 0 || matB < 0 || matA >= kMaxMaterials || matB >= kMaxMaterials) return GRAINS_ARGUMENT_ERROR;
  AdjointData* ad = s->adj;
  ACHECK(s, hipSetDevice(s->device));
  const int N = s->nparticles;
  const size_t n = (size_t)N;
  AdjArrays& A = ad->A;
  const size_t nc1 = (size_t)s->grid.ncells + 1;
  if (!ad->rec_d) {
    ACHECK(s, hipMalloc((void**)&ad->rec_d, (size_t)ad->K * rec_doubles(n) * sizeof(double)));
    ad->allocs.push_back(ad->rec_d);
    ACHECK(s, hipMalloc((void**)&ad->rec_i, (size_t)ad->K * rec_ints(n, nc1) * sizeof(int)));
    ad->allocs.push_back(ad->rec_i);
    double** slots[7] = {A.xb, A.vb, A.wb, A.Fb_in, A.Tb_in, A.Fb_out, A.Tb_out};
    for (auto sl : slots)
      for (int c = 0; c < 3; ++c) {
        ACHECK(s, hipMalloc((void**)&sl[c], n * sizeof(double)));
        ad->allocs.push_back(sl[c]);
      }
    ACHECK(s, hipMalloc((void**)&A.gth, 3 * n * sizeof(double)));
    ad->allocs.push_back(A.gth);
  }
  // storage <-> id order
  std::vector<int> ids(n);
  ACHECK(s, hipMemcpy(ids.data(), s->P.id, sizeof(int) * n, hipMemcpyDeviceToHost));
  std::vector<double> buf(n);
  auto up = [&](double* dst, const double* src, int c) -> hipError_t {
    for (int i = 0; i < N; ++i) buf[i] = src ? src[3 * (size_t)(ids[i] - 1) + c] : 0.;
    return hipMemcpy(dst, buf.data(), n * sizeof(double), hipMemcpyHostToDevice);
  };
  for (int c = 0; c < 3; ++c) {
    ACHECK(s, up(A.xb[c], dJ_dx, c));
    ACHECK(s, up(A.vb[c], dJ_dv, c));
    ACHECK(s, up(A.wb[c], dJ_dw, c));
  }
  const double h = 0.5 * s->cfg.dt, dt = s->cfg.dt;
  hipLaunchKernelGGL(k_adj_init, dim3(nblk(N, 256)), dim3(256), 0, s->stream, s->P, A, N, h);
  ACHECK(s, hipGetLastError());
  const double saved_time = s->time;
  const long saved_steps = s->steps_done;
  int e;
  for (int seg = ad->nseg - 1; seg >= 0; --seg) {
    const int n0 = seg * ad->K;
    const int len = std::min(ad->K, ad->nsteps - n0);
    // rebuild the segment from its checkpoint, recording every step
    if ((e = load_state(s, ad->ckpt[seg]))) return e;
    s->steps_done = ad->steps0 + n0;
    s->recording = true;
    s->rec_slot = 0;
    e = s->enqueue_steps(len);
    s->recording = false;
    if (e) return e;
    for (int i = len - 1; i >= 0; --i) {
      const int step = n0 + i;  // reverse of the map step -> step + 1
      // the sweep reads the recorded step in place
      double* dd = ad->rec_d + (size_t)i * rec_doubles(n);
      int* di = ad->rec_i + (size_t)i * rec_ints(n, nc1);
      Particles Pr = s->P;
      Sorted Sr = s->S;
      for (int c = 0; c < 3; ++c) {
        Pr.v[0][c] = dd + (4 + c) * n;
        Pr.w[0][c] = dd + (7 + c) * n;
      }
      Sr.posr = dd;
      Sr.order = di;
      Sr.sid = di + n;
      Sr.start = di + 2 * n;
      hipLaunchKernelGGL(k_adj_sweep, dim3(nblk(N, 64)), dim3(64), 0, s->stream, Pr, Sr, s->grid, s->d_obs,
                         (int)s->obstacles.size(), s->d_models, A, N, 0, matA, matB, h, dt, step >= 1 ? 2. : 1., 0,
                         s->d_status);
      ACHECK(s, hipGetLastError());
      for (int c = 0; c < 3; ++c) {
        std::swap(A.Fb_in[c], A.Fb_out[c]);
        std::swap(A.Tb_in[c], A.Tb_out[c]);
      }
    }
  }
  // the initial force evaluation f(x_0, v_0, w_0)
  if ((e = load_state(s, ad->ckpt[0]))) return e;
  hipLaunchKernelGGL(k_adj_sweep, dim3(nblk(N, 64)), dim3(64), 0, s->stream, s->P, s->S, s->grid, s->d_obs,
                     (int)s->obstacles.size(), s->d_models, A, N, 0, matA, matB, h, dt, 1., 1, s->d_status);
  ACHECK(s, hipGetLastError());
  ACHECK(s, hipStreamSynchronize(s->stream));
  // results
  auto down = [&](double* dst, double* src, int c) -> hipError_t {
    if (!dst) return hipSuccess;
    hipError_t er = hipMemcpy(buf.data(), src, n * sizeof(double), hipMemcpyDeviceToHost);
    for (int i = 0; i < N; ++i) dst[3 * (size_t)(ids[i] - 1) + c] = buf[i];
    return er;
  };
  for (int c = 0; c < 3; ++c) {
    ACHECK(s, down(grad_x0, A.xb[c], c));
    ACHECK(s, down(grad_v0, A.vb[c], c));
    ACHECK(s, down(grad_w0, A.wb[c], c));
  }
  if (grad_params) {
    std::vector<double> gth(3 * n);
    ACHECK(s, hipMemcpy(gth.data(), A.gth, 3 * n * sizeof(double), hipMemcpyDeviceToHost));
    double gkn = 0., gbeta = 0., gmu = 0.;
    for (size_t i = 0; i < n; ++i) {
      gkn += gth[3 * i];
      gbeta += gth[3 * i + 1];
      gmu += gth[3 * i + 2];
    }
    // beta = ln(en) / sqrt(pi^2 + ln(en)^2)  =>  dbeta/den = pi^2 / (en (pi^2 + ln^2 en)^(3/2))
    const HookeModel& hm = s->models[matA * kMaxMaterials + matB];
    double L = std::log(hm.en), q = kPi * kPi + L * L;
    grad_params[0] = gkn;
    grad_params[1] = gbeta * kPi * kPi / (hm.en * q * std::sqrt(q));
    grad_params[2] = gmu;
  }
  // back to the end of the forward run
  if ((e = load_state(s, ad->ckpt[ad->nseg]))) return e;
  s->time = saved_time;
  s->steps_done = saved_steps;
  return s->check_status();
}

}  // extern "C"

void grains_gpu_system::free_adjoint() {
  delete adj;
  adj = nullptr;
}

extern "C" int grains_contact_eval(int kind, const double* b0, const double* b1, const double* model,
                                   const double* cot, double* loads, double* grad) {
  if ((kind != 0 && kind != 1) || !b0 || !b1 || !model || !loads) return GRAINS_ARGUMENT_ERROR;
  for (int i = 0; i < 12; ++i) loads[i] = 0.;
  if (grad)
    for (int i = 0; i < 12; ++i) grad[i] = 0.;
  const double en = model[1];
  if (!(model[0] > 0.) || !(en > 0. && en < 1.)) return GRAINS_ARGUMENT_ERROR;
  const double L = std::log(en), q = kPi * kPi + L * L;
  ContactIn in;
  in.r0 = b0[3];
  in.v0 = d3{b0[4], b0[5], b0[6]};
  in.w0 = d3{b0[7], b0[8], b0[9]};
  in.m0 = b0[10];
  in.kn = model[0];
  in.beta = L / std::sqrt(q);
  in.etat = model[2];
  in.mu = model[3];
  const d3 x0{b0[0], b0[1], b0[2]};
  d3 F0b{0., 0., 0.}, T0b{0., 0., 0.}, F1b{0., 0., 0.}, T1b{0., 0., 0.};
  if (cot) {
    F0b = d3{cot[0], cot[1], cot[2]};
    T0b = d3{cot[3], cot[4], cot[5]};
    F1b = d3{cot[6], cot[7], cot[8]};
    T1b = d3{cot[9], cot[10], cot[11]};
  }
  PairGrad g{};
  if (kind == 0) {
    const d3 x1{b1[0], b1[1], b1[2]};
    in.r1 = b1[3];
    in.v1 = d3{b1[4], b1[5], b1[6]};
    in.w1 = d3{b1[7], b1[8], b1[9]};
    in.m1 = b1[10];
    const d3 AB = x1 - x0;
    const double nab = norm(AB);
    if (!(nab - (in.r0 + in.r1) < 0.) || !(nab > 0.)) return 0;
    in.u = AB / nab;
    in.d = nab - (in.r0 + in.r1);
    in.kappa = 0.5;
    in.sphere1 = true;
    if (cot && grad) pair_vjp(AB, in, F0b, T0b, F1b, T1b, g);
  } else {
    double B[9], Binv[9], io[3], ext[3];
    for (int i = 0; i < 9; ++i) B[i] = b1[6 + i];
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) Binv[3 * i + j] = B[3 * j + i];
    for (int i = 0; i < 3; ++i) {
      io[i] = -(Binv[3 * i] * b1[3] + Binv[3 * i + 1] * b1[4] + Binv[3 * i + 2] * b1[5]);
      ext[i] = 0.5 * b1[i];
    }
    in.r1 = 0.;
    in.v1 = in.w1 = d3{0., 0., 0.};
    in.m1 = kObstacleMass;
    // geometry as in wall_vjp
    double c[3], dl[3];
    int nout = 0;
    for (int a = 0; a < 3; ++a) {
      c[a] = Binv[3 * a] * x0.x + Binv[3 * a + 1] * x0.y + Binv[3 * a + 2] * x0.z + io[a];
      double e = ext[a], qc = c[a] > e ? e : (c[a] < -e ? -e : c[a]);
      nout += (c[a] > e || c[a] < -e) ? 1 : 0;
      dl[a] = c[a] - qc;
    }
    if (nout == 0) return GRAINS_ARGUMENT_ERROR;  // centre inside the box
    d3 D{B[0] * dl[0] + B[1] * dl[1] + B[2] * dl[2], B[3] * dl[0] + B[4] * dl[1] + B[5] * dl[2],
         B[6] * dl[0] + B[7] * dl[1] + B[8] * dl[2]};
    const double nd = norm(D);
    if (!(nd - in.r0 < 0.)) return 0;
    in.u = (-1. / nd) * D;
    in.d = nd - in.r0;
    in.kappa = nout <= 1 ? 0.5 : 1.5;
    in.sphere1 = false;
    if (cot && grad) wall_vjp(B, Binv, io, ext, x0, in, F0b, T0b, g);
  }
  ContactFwd w;
  contact_forward(in, w);
  const d3 T0 = cross(w.l0, w.f);
  const d3 F1 = -w.f;
  const d3 T1 = in.sphere1 ? cross(w.l1, F1) : d3{0., 0., 0.};
  const d3 out[4] = {w.f, T0, in.sphere1 ? F1 : d3{0., 0., 0.}, T1};
  for (int i = 0; i < 4; ++i) {
    loads[3 * i] = out[i].x;
    loads[3 * i + 1] = out[i].y;
    loads[3 * i + 2] = out[i].z;
  }
  if (cot && grad) {
    const d3 gs[3] = {g.x0, g.v0, g.w0};
    for (int i = 0; i < 3; ++i) {
      grad[3 * i] = gs[i].x;
      grad[3 * i + 1] = gs[i].y;
      grad[3 * i + 2] = gs[i].z;
    }
    grad[9] = g.kn;
    grad[10] = g.beta * kPi * kPi / (en * q * std::sqrt(q));
    grad[11] = g.mu;
  }
  return 1;
}
