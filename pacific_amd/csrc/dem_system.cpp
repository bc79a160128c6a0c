// Host side of the MI355X DEM time step: system construction, device memory,
// the per-step launch sequence and the C ABI declared in include/grains_gpu.h.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/grains_gpu.h"
#include "dem_device.hpp"
#include "dem_launch.hpp"
#include "dem_math.hpp"
#include "dem_shapes.hpp"
#include "dem_system.hpp"

using namespace dem;

namespace {

int fail(grains_gpu_system* s, int code, const std::string& msg) {
  if (s) s->last_error = msg;
  return code;
}

#define HIPCHECK(sys, expr)                                                                         \
  do {                                                                                              \
    hipError_t e_ = (expr);                                                                         \
    if (e_ != hipSuccess)                                                                           \
      return fail(sys, GRAINS_DEVICE_ERROR, std::string(#expr) + ": " + hipGetErrorString(e_));     \
  } while (0)

template <typename T>
hipError_t dalloc(T** p, size_t n) {
  return hipMalloc((void**)p, std::max<size_t>(n, 1) * sizeof(T));
}

// HookeContactForceModel constructor (HookeContactForceModel.cpp:11-20)
double hooke_beta(double en) { return std::log(en) / std::sqrt(kPi * kPi + std::log(en) * std::log(en)); }

}  // namespace

// ---------------------------------------------------------------------------
// Device allocation

void grains_gpu_system::free_device() {
  drop_graphs();
  for (void* p : allocations) (void)hipFree(p);
  allocations.clear();
  if (ev0) (void)hipEventDestroy(ev0);
  if (ev1) (void)hipEventDestroy(ev1);
  if (stream) (void)hipStreamDestroy(stream);
  ev0 = ev1 = nullptr;
  stream = nullptr;
}

grains_gpu_system::~grains_gpu_system() {
  if (device >= 0) (void)hipSetDevice(device);
  free_adjoint();
  free_device();
}

template <typename T>
static hipError_t track(grains_gpu_system* s, T** p, size_t n) {
  hipError_t e = dalloc(p, n);
  if (e == hipSuccess) s->allocations.push_back((void*)*p);
  return e;
}

int grains_gpu_system::allocate() {
  const size_t n = (size_t)nparticles;
  for (int b = 0; b < 2; ++b)
    for (int a = 0; a < 3; ++a) {
      HIPCHECK(this, track(this, &P.x[b][a], n));
      HIPCHECK(this, track(this, &P.v[b][a], n));
      HIPCHECK(this, track(this, &P.w[b][a], n));
    }
  for (int a = 0; a < 3; ++a) {
    HIPCHECK(this, track(this, &P.wb[a], n));
    HIPCHECK(this, track(this, &P.F[a], n));
    HIPCHECK(this, track(this, &P.T[a], n));
    if (cfg.scheme == GRAINS_ADAMS_BASHFORTH2) {
      HIPCHECK(this, track(this, &P.v_nm2[a], n));
      HIPCHECK(this, track(this, &P.dUdt_nm2[a], n));
      HIPCHECK(this, track(this, &P.wb_nm2[a], n));
      HIPCHECK(this, track(this, &P.dWdt_nm2[a], n));
    } else {
      P.v_nm2[a] = P.dUdt_nm2[a] = P.wb_nm2[a] = P.dWdt_nm2[a] = nullptr;
    }
  }
  for (int a = 0; a < 4; ++a) HIPCHECK(this, track(this, &P.q[a], n));
  HIPCHECK(this, track(this, &P.rad, n));
  HIPCHECK(this, track(this, &P.mass, n));
  HIPCHECK(this, track(this, &P.inertia, n));
  HIPCHECK(this, track(this, &P.crust, n));
  HIPCHECK(this, track(this, &P.wfac, n));
  HIPCHECK(this, track(this, &P.cfac, n));
  HIPCHECK(this, track(this, &P.req, n));
  HIPCHECK(this, track(this, &P.id, n));
  HIPCHECK(this, track(this, &P.mat, n));

  HIPCHECK(this, track(this, &S.key, n));
  HIPCHECK(this, track(this, &S.slot, n));
  HIPCHECK(this, track(this, &S.count, (size_t)grid.ncells));
  HIPCHECK(this, track(this, &S.start, (size_t)grid.ncells + 1));
  HIPCHECK(this, track(this, &S.block_sums, (size_t)scan_blocks(grid.ncells) + 1));
  HIPCHECK(this, track(this, &S.tmp_order, n));
  HIPCHECK(this, track(this, &S.order, n));
  HIPCHECK(this, track(this, &S.sid, n));
  HIPCHECK(this, track(this, &S.inv, n));
  HIPCHECK(this, track(this, &S.posr, 4 * n));

  if (has_shapes()) {
    HIPCHECK(this, track(this, &P.shape, n));
    for (int b = 0; b < 2; ++b)
      for (int j = 0; j < 9; ++j) HIPCHECK(this, track(this, &P.B[b][j], n));
    for (int a = 0; a < 3; ++a) HIPCHECK(this, track(this, &P.I3[a], n));
    HIPCHECK(this, track(this, &d_shapes, shapes.size()));
    HIPCHECK(this, track(this, &d_verts, shape_verts.size()));
    P.shapes = d_shapes;
    P.verts = d_verts;
    P.gjk = gjk_opts;
  }
  HIPCHECK(this, track(this, &d_obs, obstacles.size()));
  HIPCHECK(this, track(this, &d_models, (size_t)kMaxMaterials * kMaxMaterials));
  HIPCHECK(this, track(this, &d_status, 1));
  if (verlet)
    for (int a = 0; a < 3; ++a) HIPCHECK(this, track(this, &V.xb[a], n));
  mem = dem::ContactMemory{};
  if (has_memory() && n > 0) {
    const size_t m = (size_t)kMemoryCap * n;
    for (int b = 0; b < 2; ++b) {
      HIPCHECK(this, track(this, &mem.n[b], n));
      HIPCHECK(this, hipMemsetAsync(mem.n[b], 0, sizeof(int) * n, stream));
      HIPCHECK(this, track(this, &mem.key[b], m));
      for (int a = 0; a < 3; ++a) {
        HIPCHECK(this, track(this, &mem.kd[b][a], m));
        HIPCHECK(this, track(this, &mem.pn[b][a], m));
        HIPCHECK(this, track(this, &mem.sr[b][a], m));
      }
    }
    mem.cap = kMemoryCap;
    mem.cur = 0;
  }
  if (P.shape) {
    // a record slot per listed candidate plus one per obstacle (up to 16)
    // for the slot layout of the narrow phase, and at least kFeatCap
    const size_t fcap = std::max<size_t>(kFeatCap, (size_t)std::max(nbr_cap, 0) +
                                                       std::min<size_t>(obstacles.size(), 16));
    HIPCHECK(this, track(this, &cf.count, n));
    HIPCHECK(this, track(this, &cf.ent, fcap * n));
    HIPCHECK(this, track(this, &cf.pid, fcap * n));
    for (int j = 0; j < 9; ++j) HIPCHECK(this, track(this, &cf.f[j], fcap * n));
    HIPCHECK(this, hipMemsetAsync(cf.count, 0, sizeof(int) * n, stream));
    cf.cap = (int)fcap;
    cf.n = (int)n;
    HIPCHECK(this, track(this, &cf.items, fcap * n));
    HIPCHECK(this, track(this, &cf.nitems, 1));
    int ncu = 0;
    HIPCHECK(this, hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, device));
    cf.item_blocks = 8 * std::max(ncu, 1);  // one-wave workgroups, two per SIMD
  }
  HIPCHECK(this, track(this, &d_stat_part, 5 * ((n + 255) / 256) + 5));
  HIPCHECK(this, track(this, &d_stat_tot, 8));
  HIPCHECK(this, hipMemsetAsync(d_stat_tot, 0, 8 * sizeof(double), stream));
  if (verlet) {
    HIPCHECK(this, track(this, &d_ct_count, n));
    HIPCHECK(this, track(this, &d_ct, (size_t)kContactCap * n));
  }
  // slots of the candidate list are stored in 27 bits
  if (nbr_cap > 0 && n > 0 && n < ((size_t)1 << dem::kNbrSlotBitsHost)) {
    HIPCHECK(this, track(this, &d_nbr_count, n));
    HIPCHECK(this, track(this, &d_nbr, (size_t)nbr_cap * n));
  } else {
    d_nbr_count = nullptr;
    d_nbr = nullptr;
    verlet = false;
  }
  pair_cap = (int)std::min<size_t>(16 * n + 4096, (size_t)1 << 28);
  HIPCHECK(this, track(this, &d_pairs, 2 * (size_t)pair_cap));
  // index arrays start zeroed so that every stored index is in range
  HIPCHECK(this, hipMemsetAsync(S.tmp_order, 0, sizeof(int) * n, stream));
  HIPCHECK(this, hipMemsetAsync(S.order, 0, sizeof(int) * n, stream));
  HIPCHECK(this, hipMemsetAsync(S.sid, 0, sizeof(int) * n, stream));
  HIPCHECK(this, hipMemsetAsync(S.inv, 0, sizeof(int) * n, stream));
  HIPCHECK(this, hipMemsetAsync(S.posr, 0, sizeof(double) * 4 * n, stream));
  HIPCHECK(this, hipMemsetAsync(S.start, 0, sizeof(int) * ((size_t)grid.ncells + 1), stream));
  HIPCHECK(this, hipMemsetAsync(S.count, 0, sizeof(int) * (size_t)grid.ncells, stream));
  HIPCHECK(this, hipMemsetAsync(d_status, 0, sizeof(DevStatus), stream));
  return GRAINS_OK;
}

// ---------------------------------------------------------------------------
// Grid: LinkedCell::set (LinkedCell.cpp:47-140) with the cell size of
// Grains::Construction (Grains.cpp:1010-1022) and defineLinkedCell (:2522-2527)

int grains_gpu_system::build_grid() {
  double rmax = 0.;
  for (double r : h_rad) rmax = std::max(rmax, r);
  if (rmax < 1.e-12) rmax = 1.e16;
  double factor = cfg.cell_size_factor < 1. ? 1. : cfg.cell_size_factor;
  double cellsize = 2. * (factor * rmax);
  long total = 1;
  for (int a = 0; a < 3; ++a) {
    double L = cfg.max[a] - cfg.origin[a];
    if (!(L > 0.)) return fail(this, GRAINS_CONFIG_ERROR, "domain size must be positive on every axis");
    int n = 1;
    if (cellsize > kEps) {
      n = (int)((L + kEps) / cellsize);
      if (!n) n = 1;
    }
    grid.edge[a] = L / n;
    grid.L[a] = L;
    grid.periodic[a] = cfg.periodic[a] ? 1 : 0;
    grid.n[a] = n;
    grid.lo[a] = cfg.origin[a];
    grid.off[a] = 0;
    if (grid.periodic[a]) {
      if (n < 2)
        return fail(this, GRAINS_CONFIG_ERROR,
                    "a periodic direction needs at least 2 cells of size 2*Rmax*CellSizeFactor");
      grid.lo[a] = grid.lo[a] + (-grid.edge[a]);  // Point3::Move(-edge)
      grid.off[a] = 1;
    }
    total *= n;
  }
  if (total > (1L << 30)) return fail(this, GRAINS_CONFIG_ERROR, "too many linked cells");
  grid.ncells = (int)total;

  // Verlet list grid: cells at least 2 Rmax factor + skin wide over the
  // domain itself (no ghost layers: the list sees periodic neighbours at
  // their nearest image), three or more per periodic axis so that no cell
  // is a neighbour twice.  Spheres only.
  verlet = false;
  const double skin = verlet_skin < 0. ? 0.1 * cellsize : verlet_skin;
  if (nbr_cap > 0 && skin > 0. && !has_shapes() && cellsize > kEps && cellsize < 1.e15) {
    dem::Grid g{};
    bool ok = true;
    long tv = 1;
    for (int a = 0; a < 3; ++a) {
      double L = cfg.max[a] - cfg.origin[a];
      int n = (int)((L + kEps) / (cellsize + skin));
      if (n < 1) n = 1;
      if (grid.periodic[a] && n < 3) ok = false;
      g.edge[a] = L / n;
      g.L[a] = L;
      g.periodic[a] = grid.periodic[a];
      g.n[a] = n;
      g.lo[a] = cfg.origin[a];
      g.off[a] = 0;
      tv *= n;
    }
    g.ncells = (int)tv;
    if (ok && tv <= grid.ncells) {
      verlet = true;
      V.Gv = g;
      V.skin = skin;
      V.half2 = 0.25 * skin * skin;
    }
  }
  return GRAINS_OK;
}

// ---------------------------------------------------------------------------

int grains_gpu_system::upload() {
  const int n = nparticles;
  std::vector<double> buf((size_t)n);
  auto up = [&](double* dst, const std::vector<double>& src, int stride, int comp) -> hipError_t {
    for (int i = 0; i < n; ++i) buf[i] = src[(size_t)stride * i + comp];
    return hipMemcpy(dst, buf.data(), sizeof(double) * n, hipMemcpyHostToDevice);
  };
  for (int a = 0; a < 3; ++a) {
    HIPCHECK(this, up(P.x[0][a], h_x, 3, a));
    HIPCHECK(this, up(P.v[0][a], h_v, 3, a));
    HIPCHECK(this, up(P.w[0][a], h_w, 3, a));
    HIPCHECK(this, up(P.wb[a], h_wb, 3, a));
    HIPCHECK(this, hipMemset(P.F[a], 0, sizeof(double) * n));
    HIPCHECK(this, hipMemset(P.T[a], 0, sizeof(double) * n));
    if (P.v_nm2[a]) {
      HIPCHECK(this, hipMemset(P.v_nm2[a], 0, sizeof(double) * n));
      HIPCHECK(this, hipMemset(P.dUdt_nm2[a], 0, sizeof(double) * n));
      HIPCHECK(this, hipMemset(P.wb_nm2[a], 0, sizeof(double) * n));
      HIPCHECK(this, hipMemset(P.dWdt_nm2[a], 0, sizeof(double) * n));
    }
  }
  for (int a = 0; a < 4; ++a) HIPCHECK(this, up(P.q[a], h_q, 4, a));
  HIPCHECK(this, hipMemcpy(P.rad, h_rad.data(), sizeof(double) * n, hipMemcpyHostToDevice));
  HIPCHECK(this, hipMemcpy(P.mass, h_mass.data(), sizeof(double) * n, hipMemcpyHostToDevice));
  HIPCHECK(this, hipMemcpy(P.inertia, h_inertia.data(), sizeof(double) * n, hipMemcpyHostToDevice));
  HIPCHECK(this, hipMemcpy(P.crust, h_crust.data(), sizeof(double) * n, hipMemcpyHostToDevice));
  HIPCHECK(this, hipMemcpy(P.wfac, h_wfac.data(), sizeof(double) * n, hipMemcpyHostToDevice));
  HIPCHECK(this, hipMemcpy(P.cfac, h_cfac.data(), sizeof(double) * n, hipMemcpyHostToDevice));
  HIPCHECK(this, hipMemcpy(P.req, h_req.data(), sizeof(double) * n, hipMemcpyHostToDevice));
  std::vector<int> ids((size_t)n);
  for (int i = 0; i < n; ++i) ids[i] = i + 1;
  HIPCHECK(this, hipMemcpy(P.id, ids.data(), sizeof(int) * n, hipMemcpyHostToDevice));
  HIPCHECK(this, hipMemcpy(P.mat, h_mat.data(), sizeof(int) * n, hipMemcpyHostToDevice));
  if (!obstacles.empty())
    HIPCHECK(this, hipMemcpy(d_obs, obstacles.data(), sizeof(BoxObstacle) * obstacles.size(),
                             hipMemcpyHostToDevice));
  HIPCHECK(this, hipMemcpy(d_models, models, sizeof(models), hipMemcpyHostToDevice));
  if (has_shapes()) {
    HIPCHECK(this, hipMemcpy(P.shape, h_shape.data(), sizeof(int) * n, hipMemcpyHostToDevice));
    for (int j = 0; j < 9; ++j) {
      HIPCHECK(this, up(P.B[0][j], h_B, 9, j));
      HIPCHECK(this, up(P.B[1][j], h_B, 9, j));
    }
    for (int a = 0; a < 3; ++a) HIPCHECK(this, up(P.I3[a], h_I3, 3, a));
    HIPCHECK(this, hipMemcpy(d_shapes, shapes.data(), sizeof(ShapeDesc) * shapes.size(), hipMemcpyHostToDevice));
    if (!shape_verts.empty())
      HIPCHECK(this, hipMemcpy(d_verts, shape_verts.data(), sizeof(double) * shape_verts.size(),
                               hipMemcpyHostToDevice));
  }
  cur = 0;
  return GRAINS_OK;
}

StepParams grains_gpu_system::params() const {
  StepParams sp;
  for (int a = 0; a < 3; ++a) sp.g[a] = cfg.gravity[a];
  sp.dt = cfg.dt;
  sp.n = nparticles;
  sp.cur = cur;
  sp.nobs = (int)obstacles.size();
  sp.nmat = kMaxMaterials;
  sp.record_pairs = record_pairs;
  sp.pair_cap = pair_cap;
  sp.scheme = cfg.scheme;
  sp.first_step = 0;
  // the adjoint's replays are not counted again
  sp.stats = stats_on && !recording ? d_stat_part : nullptr;
  sp.mem = mem;
  sp.cf = cf;
  sp.fsum = fsum_on ? fsum_buf : nullptr;
  sp.hyd = d_hyd;
  return sp;
}

int grains_gpu_system::bin_current() {
  StepParams sp = params();
  if (verlet) {
    // periodic wrap of the current positions, then a Verlet list build
    HIPCHECK(this, launch_key(P, S, grid, sp, d_status, vstate(-1), stream));
    HIPCHECK(this, launch_verlet_build(P, S, V, nparticles, cur & 1, d_status, nbr_list(), 0, stream));
    return GRAINS_OK;
  }
  HIPCHECK(this, launch_key(P, S, grid, sp, d_status, vstate(0), stream));
  HIPCHECK(this, launch_bin(P, S, grid, nparticles, cur & 1, stream));
  return GRAINS_OK;
}

int grains_gpu_system::reset_pair_counter() {
  if (!record_pairs) return GRAINS_OK;
  // npairs, pair_overflow and n_slip are consecutive in DevStatus
  HIPCHECK(this, launch_clear(&d_status->npairs, 3, stream));
  return GRAINS_OK;
}

int grains_gpu_system::mark(std::vector<hipEvent_t>& evs, size_t& n) {
  if (n >= evs.size()) return fail(this, GRAINS_ARGUMENT_ERROR, "profile event buffer exhausted");
  HIPCHECK(this, hipEventRecord(evs[n++], stream));
  return GRAINS_OK;
}

// Binning after a drift; with a Verlet list, its rebuild, which runs only
// if the drift found the list stale.
int grains_gpu_system::bin_launch() {
  int e;
  if (prof && (e = mark(prof->bin_ev, prof->nbin))) return e;
  if (verlet)
    HIPCHECK(this, launch_verlet_build(P, S, V, nparticles, cur & 1, d_status, nbr_list(), kStale, stream));
  else
    HIPCHECK(this, launch_bin(P, S, grid, nparticles, cur & 1, stream));
  if (prof && (e = mark(prof->bin_ev, prof->nbin))) return e;
  return GRAINS_OK;
}

bool grains_gpu_system::has_memory() const {
  if (has_default && default_model.memory) return true;
  for (const auto& m : models)
    if (m.defined && m.memory) return true;
  return false;
}

int grains_gpu_system::forces(int mode, int first_step, bool commit) {
  int e = reset_pair_counter();
  if (e) return e;
  StepParams sp = params();
  sp.first_step = first_step;
  if (!commit) sp.stats = nullptr;
  const dem::NbrList L = nbr_list();
  if (L.count && !verlet) {
    if (prof && (e = mark(prof->detect_ev, prof->ndetect))) return e;
    HIPCHECK(this, launch_detect(S, grid, nparticles, L, d_status, stream, P.shape ? cf.ent : nullptr));
    if (prof && (e = mark(prof->detect_ev, prof->ndetect))) return e;
  }
  if (P.shape) {
    // non-spherical bodies: the narrow phase records the contacts the force
    // sweep reads
    if (prof && (e = mark(prof->narrow_ev, prof->nnarrow))) return e;
    HIPCHECK(this, launch_narrow(P, S, grid, d_obs, sp, d_status, L, stream));
    if (prof && (e = mark(prof->narrow_ev, prof->nnarrow))) return e;
  }
  const dem::NbrList CL = contact_list();
  if (CL.count) {
    // this step's contacts from the Verlet list
    if (prof && (e = mark(prof->detect_ev, prof->ndetect))) return e;
    HIPCHECK(this, launch_vfilter(P, S, grid, V, nparticles, cur & 1, L, CL, d_status, stream));
    if (prof && (e = mark(prof->detect_ev, prof->ndetect))) return e;
  }
  if (prof && (e = mark(prof->force_ev, prof->nforce))) return e;
  HIPCHECK(this, launch_forces(mode, cfg.precision == GRAINS_MIXED, P, S, grid, d_obs, d_models, sp, d_status,
                               d_pairs, L, vstate(next_stamp()), CL, stream));
  if (prof && (e = mark(prof->force_ev, prof->nforce))) return e;
  if (sp.stats) HIPCHECK(this, launch_stats_fold(d_stat_part, nparticles, d_stat_tot, force_time, stream));
  // the evaluation wrote the other memory buffer: it becomes current
  if (commit && mem.cap) mem.cur ^= 1;
  if (mode == kModeForcesKick) cur ^= 2;
  if (mode == kModeFused) cur ^= 3;
  return GRAINS_OK;
}

// After an interrupted launch sequence the per-cell counts may hold stale
// arrivals: clear them and bin the current positions again.
int grains_gpu_system::recover() {
  if (!needs_rebin) return GRAINS_OK;
  HIPCHECK(this, launch_clear(S.count, grid.ncells, stream));
  needs_rebin = false;
  return bin_current();
}

int grains_gpu_system::enqueue_steps(int nsteps) {
  int e = recover();
  if (e) return e;
  e = enqueue_steps_impl(nsteps);
  if (e) needs_rebin = true;
  return e;
}

// One batch of time steps (Grains::Simulation loop body, Grains.cpp:286-452)
int grains_gpu_system::enqueue_steps_impl(int nsteps) {
  if (nsteps <= 0) return GRAINS_OK;
  int e;
  if (cfg.scheme == GRAINS_LEAPFROG) {
    // Move(dt/2, dt) -> LinkUpdate -> forces -> advanceVelocity(dt/2); the
    // forces, the second half kick and the next Move share one kernel.
    StepParams sp = params();
    // forces of step j are evaluated at the time the step reaches (the
    // reference adds dt to m_time before them, Grains.cpp:290)
    force_time = time;
    HIPCHECK(this, launch_move(P, S, grid, sp, d_status, 0.5 * cfg.dt, cfg.dt, vstate(next_stamp()), stream));
    if ((e = bin_launch())) return e;
    if (recording && (e = record_state())) return e;
    int i = 1;
    if (graph_steps_allowed())
      for (bool done = true; done && nsteps - i >= kGraphSteps; i += done ? kGraphSteps : 0)
        if ((e = graph_chunk(done))) return e;
    for (; i < nsteps; ++i) {
      force_time += cfg.dt;
      if ((e = forces(kModeFused, 0))) return e;
      if ((e = bin_launch())) return e;
      if (recording && (e = record_state())) return e;
    }
    force_time += cfg.dt;
    if ((e = forces(kModeForcesKick, 0))) return e;
  } else {
    // forces -> Move(dt, dt) -> LinkUpdate (Grains.cpp:365-377)
    force_time = time;
    for (int i = 0; i < nsteps; ++i) {
      int first = (cfg.scheme == GRAINS_ADAMS_BASHFORTH2 && steps_done + i == 0) ? 1 : 0;
      force_time += cfg.dt;
      if ((e = forces(kModeFused, first))) return e;
      if ((e = bin_launch())) return e;
    }
  }
  steps_done += nsteps;
  for (int i = 0; i < nsteps; ++i) time += cfg.dt;
  return GRAINS_OK;
}

// Graph replay serves the plain time loop: no per-step host work (profile
// events, contact records, statistics with their time stamps, adjoint
// recording) may be part of it.
// PACIFIC_AMD_NO_GRAPHS=1 in the environment launches every step directly
// (profiling runs: per-kernel tracers see the same kernels either way).
bool grains_gpu_system::graph_steps_allowed() const {
  static const bool disabled = [] {
    const char* v = std::getenv("PACIFIC_AMD_NO_GRAPHS");
    return v && v[0] == '1';
  }();
  return !disabled && !prof && !recording && !record_pairs && !stats_on && cfg.scheme == GRAINS_LEAPFROG;
}

void grains_gpu_system::drop_graphs() {
  for (auto& g : graphs) (void)hipGraphExecDestroy(g.second);
  graphs.clear();
}

// kGraphSteps fused LeapFrog steps (forces, kicks, drift, binning or the
// gated Verlet rebuild): captured from the launches of the plain loop the
// first time a parity state is met, replayed afterwards.  The kernels read
// every per-step quantity from device memory, so the captured arguments
// hold for every replay that starts from the same (cur, mem.cur).
// done = false (and no error): the steps could not be captured on this
// stack; nothing was launched, the host state is as before, and the plain
// loop carries on.
int grains_gpu_system::graph_chunk(bool& done) {
  done = false;
  if (graphs_failed) return GRAINS_OK;
  if (cfg.dt != graph_dt) {
    drop_graphs();
    graph_dt = cfg.dt;
  }
  const int key = (cur & 3) | ((mem.cur & 1) << 2);
  for (auto& g : graphs)
    if (g.first == key) {
      for (int j = 0; j < kGraphSteps; ++j) force_time += cfg.dt;
      HIPCHECK(this, hipGraphLaunch(g.second, stream));
      ++graph_replays;
      done = true;
      return GRAINS_OK;
    }
  const int cur0 = cur, mem0 = mem.cur;
  const double ft0 = force_time;
  HIPCHECK(this, hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
  int e = GRAINS_OK;
  for (int j = 0; j < kGraphSteps && !e; ++j) {
    force_time += cfg.dt;
    e = forces(kModeFused, 0);
    if (!e) e = bin_launch();
  }
  hipGraph_t graph = nullptr;
  const hipError_t ce = hipStreamEndCapture(stream, &graph);
  hipGraphExec_t exec = nullptr;
  hipError_t ie = hipErrorUnknown;
  if (!e && ce == hipSuccess && graph && cur == cur0 && mem.cur == mem0)
    ie = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
  if (graph) (void)hipGraphDestroy(graph);
  if (ie != hipSuccess) {
    graph_error = e ? last_error : std::string(hipGetErrorString(ce != hipSuccess ? ce : ie));
    // captured launches never ran: back to the state before the capture
    cur = cur0;
    mem.cur = mem0;
    force_time = ft0;
    graphs_failed = true;
    (void)hipGetLastError();
    return GRAINS_OK;
  }
  graphs.emplace_back(key, exec);
  HIPCHECK(this, hipGraphLaunch(exec, stream));
  ++graph_replays;
  done = true;
  return GRAINS_OK;
}

int grains_gpu_system::check_status() {
  DevStatus st;
  HIPCHECK(this, hipMemcpyAsync(&st, d_status, sizeof(DevStatus), hipMemcpyDeviceToHost, stream));
  HIPCHECK(this, hipStreamSynchronize(stream));
  last_status = st;
  if (st.code) {
    static const char* names[] = {"", "ContactError", "MotionError", "SimulationError"};
    char buf[160];
    if (st.code == kMemoryOverflow) {
      snprintf(buf, sizeof(buf), "particle %d has %d contacts with memory, more than the %d kept per particle", st.id0,
               st.id1, kMemoryCap);
      return fail(this, GRAINS_DEVICE_ERROR, buf);
    }
    if (st.code == kContactOverflow) {
      snprintf(buf, sizeof(buf), "particle %d has %d contacts, more than the %d recorded per particle", st.id0,
               st.id1, cf.cap);
      return fail(this, GRAINS_SIMULATION_ERROR, buf);
    }
    snprintf(buf, sizeof(buf), "%s raised on the device (bodies %d and %d)", names[st.code & 3], st.id0, st.id1);
    return fail(this, st.code, buf);
  }
  if (st.pair_overflow) return fail(this, GRAINS_DEVICE_ERROR, "contact record buffer overflow");
  return GRAINS_OK;
}

int grains_gpu_system::initialize() {
  if (initialized) return fail(this, GRAINS_ARGUMENT_ERROR, "system already initialised");
  if (cfg.scheme < 0 || cfg.scheme > 3) return fail(this, GRAINS_CONFIG_ERROR, "unknown time integrator");
  if (!(cfg.dt > 0.)) return fail(this, GRAINS_CONFIG_ERROR, "time step must be positive");
  // ContactBuilderFactory::checkContactForceModelsExist (ContactBuilderFactory.cpp:31-76):
  // every material pair involving a particle material needs a model
  std::vector<int> used(kMaxMaterials, 0);
  for (int m : h_mat) used[m] |= 1;
  for (auto& o : obstacles) used[o.mat] |= 2;
  for (int a = 0; a < kMaxMaterials; ++a)
    for (int b = 0; b < kMaxMaterials; ++b)
      if ((used[a] & 1) && used[b] && !models[a * kMaxMaterials + b].defined) {
        if (has_default) models[a * kMaxMaterials + b] = default_model;
        else {
          char buf[128];
          snprintf(buf, sizeof(buf), "no contact force model for materials %d and %d", a, b);
          return fail(this, GRAINS_CONFIG_ERROR, buf);
        }
      }
  if (has_shapes() && cfg.scheme != GRAINS_LEAPFROG)
    return fail(this, GRAINS_CONFIG_ERROR, "non-spherical particles are integrated with SecondOrderLeapFrog only");
  int e = build_grid();
  if (e) return e;
  HIPCHECK(this, hipSetDevice(device));
  HIPCHECK(this, hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
  HIPCHECK(this, hipEventCreate(&ev0));
  HIPCHECK(this, hipEventCreate(&ev1));
  if ((e = allocate())) return e;
  if ((e = upload())) return e;
  if ((e = bin_current())) return e;
  initialized = true;
  // LeapFrog: accelerations at the initial time (Grains.cpp:169-174)
  if (cfg.scheme == GRAINS_LEAPFROG) {
    force_time = time;
    if ((e = forces(kModeForcesOnly, 0))) return e;
  }
  return check_status();
}

int grains_gpu_system::grow(const std::function<int()>& add) {
  if (has_memory())
    return fail(this, GRAINS_CONFIG_ERROR, "insertion over time is not supported with a contact model with memory");
  if (cfg.scheme == GRAINS_ADAMS_BASHFORTH2)
    return fail(this, GRAINS_CONFIG_ERROR, "insertion over time is not supported with SecondOrderAdamsBashforth");
  const size_t n = (size_t)nparticles;
  std::vector<double> F(3 * n), T(3 * n);
  h_x.resize(3 * n);
  h_v.resize(3 * n);
  h_w.resize(3 * n);
  h_wb.resize(3 * n);
  h_q.resize(4 * n);
  int e = grains_gpu_get_state(this, h_x.data(), h_v.data(), h_w.data(), h_wb.data(), h_q.data(), F.data(), T.data());
  if (e) return e;
  if (has_shapes()) {
    h_B.resize(9 * n);
    if ((e = grains_gpu_get_basis(this, h_B.data()))) return e;
  }
  double stat[8];
  DevStatus st;
  HIPCHECK(this, hipMemcpy(stat, d_stat_tot, sizeof(stat), hipMemcpyDeviceToHost));
  HIPCHECK(this, hipMemcpy(&st, d_status, sizeof(st), hipMemcpyDeviceToHost));
  const double ft = force_time;
  free_adjoint();
  free_device();
  d_hyd = d_fsum = fsum_buf = nullptr;
  fsum_on = false;
  initialized = false;
  needs_rebin = false;
  cur = 0;
  if ((e = add())) return e;
  if ((e = initialize())) return e;
  // the carried forces of the particles already in; the new one keeps the
  // weight of the evaluation at initialisation (it touches nothing)
  const size_t m = (size_t)nparticles;
  std::vector<int> ids(m);
  HIPCHECK(this, hipMemcpy(ids.data(), P.id, sizeof(int) * m, hipMemcpyDeviceToHost));
  std::vector<double> buf(m);
  for (int a = 0; a < 3; ++a)
    for (int k = 0; k < 2; ++k) {
      double* dev = k ? P.T[a] : P.F[a];
      const std::vector<double>& old = k ? T : F;
      HIPCHECK(this, hipMemcpy(buf.data(), dev, sizeof(double) * m, hipMemcpyDeviceToHost));
      for (size_t i = 0; i < m; ++i)
        if ((size_t)(ids[i] - 1) < n) buf[i] = old[3 * (size_t)(ids[i] - 1) + a];
      HIPCHECK(this, hipMemcpy(dev, buf.data(), sizeof(double) * m, hipMemcpyHostToDevice));
    }
  HIPCHECK(this, hipMemcpy(d_stat_tot, stat, sizeof(stat), hipMemcpyHostToDevice));
  HIPCHECK(this, hipMemcpy(d_status, &st, sizeof(st), hipMemcpyHostToDevice));
  force_time = ft;
  return GRAINS_OK;
}

// ---------------------------------------------------------------------------
// C ABI

extern "C" {

int grains_gpu_create(const grains_gpu_config* cfg, grains_gpu_system** out) {
  if (!cfg || !out) return GRAINS_ARGUMENT_ERROR;
  grains_gpu_system* s = new grains_gpu_system();
  s->cfg = *cfg;
  s->device = cfg->device;
  s->time = cfg->t_start;
  *out = s;
  return GRAINS_OK;
}

void grains_gpu_destroy(grains_gpu_system* sys) { delete sys; }

static int set_model(grains_gpu_system* s, int matA, int matB, const HookeModel& m) {
  if (s->initialized) {
    // initialize() copied the default into every pair left undefined and
    // sized the contact memory from the models then set: a default or a
    // change of model kind now would be silently ignored
    if (matA < 0 || matB < 0)
      return fail(s, GRAINS_ARGUMENT_ERROR, "the default contact model is set before initialisation");
    if (s->models[matA * kMaxMaterials + matB].memory != m.memory)
      return fail(s, GRAINS_ARGUMENT_ERROR,
                  "a contact model with memory replaces or is replaced by one without before initialisation");
  }
  if (matA < 0 || matB < 0) {
    s->default_model = m;
    s->has_default = true;
  } else {
    s->models[matA * kMaxMaterials + matB] = m;
    s->models[matB * kMaxMaterials + matA] = m;
  }
  if (s->initialized)
    HIPCHECK(s, hipMemcpy(s->d_models, s->models, sizeof(s->models), hipMemcpyHostToDevice));
  return GRAINS_OK;
}

int grains_gpu_set_contact_model(grains_gpu_system* s, int matA, int matB, double kn, double en, double etat,
                                 double muc, double kr) {
  if (!s) return GRAINS_ARGUMENT_ERROR;
  if (matA >= kMaxMaterials || matB >= kMaxMaterials) return fail(s, GRAINS_ARGUMENT_ERROR, "material index too large");
  if (!(en > 0.) || en > 1.) return fail(s, GRAINS_CONFIG_ERROR, "en must lie in (0, 1]");
  HookeModel m{};
  m.kn = kn; m.en = en; m.etat = etat; m.muc = muc; m.kr = kr;
  m.beta = hooke_beta(en);
  m.defined = 1;
  m.memory = 0;
  return set_model(s, matA, matB, m);
}

int grains_gpu_set_memory_contact_model(grains_gpu_system* s, int matA, int matB, const double* p) {
  if (!s || !p) return GRAINS_ARGUMENT_ERROR;
  if (matA >= kMaxMaterials || matB >= kMaxMaterials) return fail(s, GRAINS_ARGUMENT_ERROR, "material index too large");
  if (!(p[1] > 0.) || p[1] > 1.) return fail(s, GRAINS_CONFIG_ERROR, "en must lie in (0, 1]");
  if (s->initialized) return fail(s, GRAINS_ARGUMENT_ERROR, "a model with memory is defined before initialisation");
  HookeModel m{};
  m.kn = p[0]; m.en = p[1]; m.etat = p[2]; m.muc = p[3]; m.kr = 0.;
  m.kt = p[4]; m.mur = p[5]; m.etarpf = p[6]; m.epst = p[7];
  m.beta = hooke_beta(m.en);
  m.defined = 1;
  m.memory = 1;
  return set_model(s, matA, matB, m);
}


int grains_gpu_add_spheres(grains_gpu_system* s, size_t n, const double* radius, const double* density,
                           const double* crust, const int* material, const double* x, const double* v,
                           const double* w) {
  if (!s || !radius || !density || !crust || !material || !x) return GRAINS_ARGUMENT_ERROR;
  if (s->initialized) return fail(s, GRAINS_ARGUMENT_ERROR, "particles must be added before initialisation");
  for (size_t i = 0; i < n; ++i) {
    double r = radius[i], rho = density[i];
    if (!(r > 0.) || !(rho > 0.) || !(crust[i] > 0.))
      return fail(s, GRAINS_CONFIG_ERROR, "radius, density and crust thickness must be positive");
    if (material[i] < 0 || material[i] >= kMaxMaterials) return fail(s, GRAINS_ARGUMENT_ERROR, "bad material");
    // Particle(DOMNode*) mass and inertia (Particle.cpp:157-162, Sphere.cpp:58-64, 152-155)
    double volume = 4.0 * kPi * std::pow(r, 3.) / 3.0;
    double mass = rho * volume;
    double I = 8.0 * kPi * std::pow(r, 5.) / 15.0;
    I *= rho;
    double buoy = 1. - s->cfg.fluid_density / rho;
    s->h_rad.push_back(r);
    s->h_mass.push_back(mass);
    s->h_inertia.push_back(I);
    s->h_crust.push_back(crust[i]);
    s->h_wfac.push_back(mass * buoy);  // Particle::computeWeight (Particle.cpp:413-418)
    double cf = 1.;
    if (s->cfg.fluid_density != 0.) cf -= s->cfg.fluid_density / rho;  // ParticleKinematics.cpp:485-487
    if (!s->fluid_corrected) cf = 1.;  // Particle::setFluidCorrectedAcceleration(false)
    s->h_cfac.push_back(cf);
    s->h_density.push_back(rho);
    s->h_req.push_back(std::pow((0.75 / kPi) * volume, 1. / 3.));  // Component.cpp:205-208
    s->h_mat.push_back(material[i]);
    d3 wv{0., 0., 0.};
    for (int a = 0; a < 3; ++a) {
      s->h_x.push_back(x[3 * i + a]);
      s->h_v.push_back(v ? v[3 * i + a] : 0.);
      s->h_w.push_back(w ? w[3 * i + a] : 0.);
    }
    if (w) wv = d3{w[3 * i], w[3 * i + 1], w[3 * i + 2]};
    quat q{d3{0., 0., 0.}, 1.};
    d3 wb = to_body(q, wv);  // ParticleKinematics::setAngularVelocity
    s->h_wb.push_back(wb.x);
    s->h_wb.push_back(wb.y);
    s->h_wb.push_back(wb.z);
    s->h_q.push_back(0.);
    s->h_q.push_back(0.);
    s->h_q.push_back(0.);
    s->h_q.push_back(1.);
    s->h_shape.push_back(-1);
    s->h_class.push_back(s->next_class);
    for (int j = 0; j < 9; ++j) s->h_B.push_back(j % 4 == 0 ? 1. : 0.);
    for (int a = 0; a < 3; ++a) s->h_I3.push_back(I);
  }
  s->nparticles = (int)s->h_rad.size();
  return GRAINS_OK;
}

int grains_gpu_add_shape(grains_gpu_system* s, int type, double radius, const double* L, int nv,
                         const double* verts, int nf, const int* face_sizes, const int* face_idx, double crust,
                         int* shape_id) {
  if (!s || !shape_id) return GRAINS_ARGUMENT_ERROR;
  if (s->initialized) return fail(s, GRAINS_ARGUMENT_ERROR, "shapes must be added before initialisation");
  ShapeBuild b;
  std::string err;
  if (build_shape(type, radius, L, nv, verts, nf, face_sizes, face_idx, crust, b, err))
    return fail(s, GRAINS_CONFIG_ERROR, err);
  b.desc.vert_off = (int)(s->shape_verts.size() / 3);
  s->shape_verts.insert(s->shape_verts.end(), b.verts.begin(), b.verts.end());
  std::vector<int> faces;
  if (type == kShapePoly) {
    int off = 0;
    for (int f = 0; f < nf; ++f) {
      faces.push_back(face_sizes[f]);
      for (int j = 0; j < face_sizes[f]; ++j) faces.push_back(face_idx[off + j]);
      off += face_sizes[f];
    }
  }
  s->shape_faces.push_back(faces);
  s->shapes.push_back(b.desc);
  s->shape_volume.push_back(b.volume);
  for (int i = 0; i < 6; ++i) s->shape_inertia.push_back(b.inertia[i]);
  *shape_id = (int)s->shapes.size() - 1;
  return GRAINS_OK;
}

int grains_gpu_add_particles(grains_gpu_system* s, int shape, size_t n, const double* density, const int* material,
                             const double* x, const double* v, const double* w, const double* basis) {
  if (!s || !density || !material || !x) return GRAINS_ARGUMENT_ERROR;
  if (s->initialized) return fail(s, GRAINS_ARGUMENT_ERROR, "particles must be added before initialisation");
  if (shape < 0 || shape >= (int)s->shapes.size()) return fail(s, GRAINS_ARGUMENT_ERROR, "unknown shape");
  const ShapeDesc& d = s->shapes[shape];
  const double vol = s->shape_volume[shape];
  const double* Iu = &s->shape_inertia[6 * (size_t)shape];
  for (size_t i = 0; i < n; ++i) {
    const double rho = density[i];
    if (!(rho > 0.)) return fail(s, GRAINS_CONFIG_ERROR, "density must be positive");
    if (material[i] < 0 || material[i] >= kMaxMaterials) return fail(s, GRAINS_ARGUMENT_ERROR, "bad material");
    // Particle(DOMNode*) (Particle.cpp:157-166): mass = rho V, inertia = rho I
    const double mass = rho * vol;
    double I[6];
    for (int k = 0; k < 6; ++k) I[k] = Iu[k] * rho;
    const double buoy = 1. - s->cfg.fluid_density / rho;
    s->h_rad.push_back(d.type == kShapeSphere ? d.radius : d.circR);
    s->h_mass.push_back(mass);
    s->h_inertia.push_back(I[0]);
    s->h_crust.push_back(d.crust);
    s->h_wfac.push_back(mass * buoy);
    double cf = 1.;
    if (s->cfg.fluid_density != 0.) cf -= s->cfg.fluid_density / rho;
    if (!s->fluid_corrected) cf = 1.;
    s->h_cfac.push_back(cf);
    s->h_density.push_back(rho);
    s->h_req.push_back(std::pow((0.75 / kPi) * vol, 1. / 3.));
    s->h_mat.push_back(material[i]);
    double Bm[9], q[4];
    for (int j = 0; j < 9; ++j) Bm[j] = basis ? basis[9 * i + j] : (j % 4 == 0 ? 1. : 0.);
    if (basis) quat_from_basis(Bm, q);
    else q[0] = q[1] = q[2] = 0., q[3] = 1.;
    quat qq{d3{q[0], q[1], q[2]}, q[3]};
    d3 wv{0., 0., 0.};
    for (int a = 0; a < 3; ++a) {
      s->h_x.push_back(x[3 * i + a]);
      s->h_v.push_back(v ? v[3 * i + a] : 0.);
      s->h_w.push_back(w ? w[3 * i + a] : 0.);
    }
    if (w) wv = d3{w[3 * i], w[3 * i + 1], w[3 * i + 2]};
    d3 wb = to_body(qq, wv);
    s->h_wb.push_back(wb.x);
    s->h_wb.push_back(wb.y);
    s->h_wb.push_back(wb.z);
    for (int a = 0; a < 4; ++a) s->h_q.push_back(q[a]);
    s->h_shape.push_back(shape);
    s->h_class.push_back(s->next_class);
    for (int j = 0; j < 9; ++j) s->h_B.push_back(Bm[j]);
    s->h_I3.push_back(I[0]);
    s->h_I3.push_back(I[3]);
    s->h_I3.push_back(I[5]);
  }
  s->nparticles = (int)s->h_rad.size();
  return GRAINS_OK;
}

int grains_gpu_set_collision(grains_gpu_system* s, int gjk_sv, double tolerance, int acceleration,
                             int bounding_volume) {
  if (!s) return GRAINS_ARGUMENT_ERROR;
  if (s->initialized) return fail(s, GRAINS_ARGUMENT_ERROR, "collision options are set before initialisation");
  if (!(tolerance >= 1e-10)) return fail(s, GRAINS_CONFIG_ERROR, "GJK tolerance should be at least 1E-10");
  if (bounding_volume < 0 || bounding_volume > 2) return fail(s, GRAINS_ARGUMENT_ERROR, "bounding volume 0, 1 or 2");
  s->gjk_opts = dem::GjkOpts{gjk_sv ? 1 : 0, acceleration ? 1 : 0, bounding_volume, 0, tolerance};
  return GRAINS_OK;
}

int grains_gpu_closest_point(grains_gpu_system* s, int shapeA, const double* xA, const double* BA, int shapeB,
                             const double* xB, const double* BB, double* out) {
  if (!s || !xA || !BA || !xB || !BB || !out) return GRAINS_ARGUMENT_ERROR;
  const int ns = (int)s->shapes.size();
  if (shapeA < 0 || shapeA >= ns || shapeB < 0 || shapeB >= ns) return fail(s, GRAINS_ARGUMENT_ERROR, "unknown shape");
  auto geo = [&](int sh, const double* x, const double* B) {
    const ShapeDesc& d = s->shapes[sh];
    ConvexGeo g;
    g.type = d.type;
    g.nv = d.nv;
    g.verts = s->shape_verts.data() + 3 * (size_t)d.vert_off;
    g.r = d.radius;
    g.e = d3{d.extent[0], d.extent[1], d.extent[2]};
    for (int j = 0; j < 9; ++j) g.B[j] = B[j];
    g.o = d3{x[0], x[1], x[2]};
    g.scaling = d3{d.scaling[0], d.scaling[1], d.scaling[2]};
    g.bext = d3{d.bext[0], d.bext[1], d.bext[2]};
    g.crust = d.crust;
    g.circR = d.circR;
    g.obb = d3{d.obb[0], d.obb[1], d.obb[2]};
    g.opt = s->gjk_opts;
    return g;
  };
  ConvexGeo A = geo(shapeA, xA, BA), B = geo(shapeB, xB, BB);
  d3 P{0., 0., 0.}, ov{0., 0., 0.};
  double dd = 1e20;
  int res;
  if (A.type == kShapeSphere && B.type == kShapeSphere) {
    d3 AB = B.o - A.o;
    double nab = norm(AB);
    res = sphere_sphere(A.o, A.r, A.crust, AB, nab, B.r, B.crust, P, ov, dd);
    if (res < 0 && recover_ss(A.o, A.r, A.crust, B.o, B.r, B.crust, P, ov, dd)) res = 2;
  } else {
    res = closest_convex(A, B, P, ov, dd);
    if (res < 0 && recover_gjk(A, B, P, ov, dd)) res = 2;
  }
  out[0] = P.x; out[1] = P.y; out[2] = P.z;
  out[3] = ov.x; out[4] = ov.y; out[5] = ov.z;
  out[6] = dd;
  out[7] = res;
  return GRAINS_OK;
}

int grains_gpu_shape_info(grains_gpu_system* s, int shape, double* out) {
  if (!s || !out || shape < 0 || shape >= (int)s->shapes.size()) return GRAINS_ARGUMENT_ERROR;
  const ShapeDesc& d = s->shapes[shape];
  out[0] = s->shape_volume[shape];
  out[1] = d.circR;
  for (int i = 0; i < 6; ++i) out[2 + i] = s->shape_inertia[6 * (size_t)shape + i];
  for (int a = 0; a < 3; ++a) out[8 + a] = d.scaling[a];
  for (int a = 0; a < 3; ++a) out[11 + a] = d.bext[a];
  return GRAINS_OK;
}

int grains_gpu_add_box(grains_gpu_system* s, const double* L, const double* centre, const double* basis, double crust,
                       int material) {
  if (!s || !L || !centre) return GRAINS_ARGUMENT_ERROR;
  if (s->initialized) return fail(s, GRAINS_ARGUMENT_ERROR, "obstacles must be added before initialisation");
  if (material < 0 || material >= kMaxMaterials) return fail(s, GRAINS_ARGUMENT_ERROR, "bad material");
  BoxObstacle o;
  std::memset(&o, 0, sizeof(o));
  o.id = -(int)s->obstacles.size() - 1;
  double e[3] = {std::fabs(L[0]) / 2., std::fabs(L[1]) / 2., std::fabs(L[2]) / 2.};
  for (int a = 0; a < 3; ++a) {
    o.extent[a] = e[a];
    o.centre[a] = centre[a];
  }
  for (int i = 0; i < 9; ++i) o.B[i] = basis ? basis[i] : ((i % 4) == 0 ? 1. : 0.);
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) o.Binv[3 * i + j] = o.B[3 * j + i];
  d3 c{centre[0], centre[1], centre[2]};
  // Transform::setToInverseTransform (Transform.cpp:71-77)
  for (int i = 0; i < 3; ++i) o.invorigin[i] = dot(-d3{o.Binv[3 * i], o.Binv[3 * i + 1], o.Binv[3 * i + 2]}, c);
  // Box::setCornersFaces (Box.cpp:92-110)
  const double cs[8][3] = {{-e[0], -e[1], -e[2]}, {e[0], -e[1], -e[2]}, {e[0], -e[1], e[2]}, {-e[0], -e[1], e[2]},
                           {-e[0], e[1], -e[2]},  {e[0], e[1], -e[2]},  {e[0], e[1], e[2]},  {-e[0], e[1], e[2]}};
  for (int k = 0; k < 8; ++k)
    for (int a = 0; a < 3; ++a) o.corners[3 * k + a] = cs[k][a];
  o.crust = crust;
  o.mat = material;
  o.req = std::pow((0.75 / kPi) * (8.0 * e[0] * e[1] * e[2]), 1. / 3.);
  // Convex::bbox with Box::support (Convex.cpp:29-41, Box.cpp:307-316), BBox::setValue (BBox.cpp:235-239)
  d3 ext{e[0], e[1], e[2]};
  auto support = [&](d3 d) {
    if (norm(d) < kEps) return d3{0., 0., 0.};
    return d3{d.x < 0. ? -ext.x : ext.x, d.y < 0. ? -ext.y : ext.y, d.z < 0. ? -ext.z : ext.z};
  };
  double mn[3], mx[3];
  for (int a = 0; a < 3; ++a) {
    d3 row{o.B[3 * a], o.B[3 * a + 1], o.B[3 * a + 2]};
    mn[a] = centre[a] + dot(row, support(-row));
    mx[a] = centre[a] + dot(row, support(row));
  }
  for (int a = 0; a < 3; ++a) {
    o.aabb_e[a] = (mx[a] - mn[a]) / 2.;
    o.aabb_c[a] = mn[a] + o.aabb_e[a];
  }
  s->obstacles.push_back(o);
  s->obs_to_fluid.push_back(0);
  return GRAINS_OK;
}

int grains_gpu_initialize(grains_gpu_system* s) {
  if (!s) return GRAINS_ARGUMENT_ERROR;
  return s->initialize();
}

int grains_gpu_step(grains_gpu_system* s, int nsteps) {
  if (!s) return GRAINS_ARGUMENT_ERROR;
  if (!s->initialized) return fail(s, GRAINS_ARGUMENT_ERROR, "system not initialised");
  HIPCHECK(s, hipSetDevice(s->device));
  int e = s->enqueue_steps(nsteps);
  if (e) return e;
  return s->check_status();
}

int grains_gpu_step_timed(grains_gpu_system* s, int nsteps, double* ms) {
  if (!s || !ms) return GRAINS_ARGUMENT_ERROR;
  if (!s->initialized) return fail(s, GRAINS_ARGUMENT_ERROR, "system not initialised");
  HIPCHECK(s, hipSetDevice(s->device));
  HIPCHECK(s, hipEventRecord(s->ev0, s->stream));
  int e = s->enqueue_steps(nsteps);
  if (e) return e;
  HIPCHECK(s, hipEventRecord(s->ev1, s->stream));
  HIPCHECK(s, hipEventSynchronize(s->ev1));
  float t = 0.f;
  HIPCHECK(s, hipEventElapsedTime(&t, s->ev0, s->ev1));
  *ms = t;
  return s->check_status();
}

int grains_gpu_profile_steps(grains_gpu_system* s, int nsteps, double* out) {
  if (!s || !out || nsteps <= 0 || nsteps > 100000) return GRAINS_ARGUMENT_ERROR;
  if (!s->initialized) return fail(s, GRAINS_ARGUMENT_ERROR, "system not initialised");
  HIPCHECK(s, hipSetDevice(s->device));
  grains_gpu_system::Profile pr;
  pr.force_ev.resize(2 * (size_t)nsteps + 2);
  pr.bin_ev.resize(2 * (size_t)nsteps + 2);
  pr.detect_ev.resize(2 * (size_t)nsteps + 2);
  pr.narrow_ev.resize(2 * (size_t)nsteps + 2);
  int e = GRAINS_OK;
  for (auto& ev : pr.force_ev) HIPCHECK(s, hipEventCreate(&ev));
  for (auto& ev : pr.bin_ev) HIPCHECK(s, hipEventCreate(&ev));
  for (auto& ev : pr.detect_ev) HIPCHECK(s, hipEventCreate(&ev));
  for (auto& ev : pr.narrow_ev) HIPCHECK(s, hipEventCreate(&ev));
  HIPCHECK(s, hipEventRecord(s->ev0, s->stream));
  s->prof = &pr;
  e = s->enqueue_steps(nsteps);
  s->prof = nullptr;
  if (!e) {
    HIPCHECK(s, hipEventRecord(s->ev1, s->stream));
    HIPCHECK(s, hipEventSynchronize(s->ev1));
    float t = 0.f;
    HIPCHECK(s, hipEventElapsedTime(&t, s->ev0, s->ev1));
    out[0] = t;
    double fsum = 0., bsum = 0.;
    for (size_t i = 0; i + 1 < pr.nforce; i += 2) {
      HIPCHECK(s, hipEventElapsedTime(&t, pr.force_ev[i], pr.force_ev[i + 1]));
      fsum += t;
    }
    for (size_t i = 0; i + 1 < pr.nbin; i += 2) {
      HIPCHECK(s, hipEventElapsedTime(&t, pr.bin_ev[i], pr.bin_ev[i + 1]));
      bsum += t;
    }
    out[1] = fsum;
    out[2] = (double)(pr.nforce / 2);
    out[3] = bsum;
    out[4] = (double)(pr.nbin / 2);
    double dsum = 0.;
    for (size_t i = 0; i + 1 < pr.ndetect; i += 2) {
      HIPCHECK(s, hipEventElapsedTime(&t, pr.detect_ev[i], pr.detect_ev[i + 1]));
      dsum += t;
    }
    out[5] = dsum;
    out[6] = (double)(pr.ndetect / 2);
    double nsum = 0.;
    for (size_t i = 0; i + 1 < pr.nnarrow; i += 2) {
      HIPCHECK(s, hipEventElapsedTime(&t, pr.narrow_ev[i], pr.narrow_ev[i + 1]));
      nsum += t;
    }
    out[7] = nsum;
    out[8] = (double)(pr.nnarrow / 2);
  }
  for (auto& ev : pr.force_ev) (void)hipEventDestroy(ev);
  for (auto& ev : pr.bin_ev) (void)hipEventDestroy(ev);
  for (auto& ev : pr.detect_ev) (void)hipEventDestroy(ev);
  for (auto& ev : pr.narrow_ev) (void)hipEventDestroy(ev);
  if (e) return e;
  return s->check_status();
}

int grains_gpu_graph_stats(grains_gpu_system* s, int* ngraphs, long* replays, char* why, size_t cap) {
  if (!s) return GRAINS_ARGUMENT_ERROR;
  if (ngraphs) *ngraphs = s->graphs_failed ? -1 : (int)s->graphs.size();
  if (replays) *replays = s->graph_replays;
  if (why && cap) {
    std::snprintf(why, cap, "%s", s->graph_error.c_str());
  }
  return GRAINS_OK;
}

int grains_gpu_synchronize(grains_gpu_system* s) {
  if (!s || !s->initialized) return GRAINS_ARGUMENT_ERROR;
  return s->check_status();
}

int grains_gpu_compute_forces(grains_gpu_system* s) {
  if (!s || !s->initialized) return GRAINS_ARGUMENT_ERROR;
  HIPCHECK(s, hipSetDevice(s->device));
  // an evaluation outside the time loop neither enters the statistics nor
  // advances the contact memory
  int e = s->forces(kModeForcesOnly, 0, false);
  if (e) return e;
  return s->check_status();
}

// GrainsCoupledWithFluid::updateParticlesHydroFT (GrainsCoupledWithFluid.cpp:
// 1187-1191) -> AppPRSHydroFT::setHydroFT: m rows of (Fx, Fy, Fz, Tx, Ty, Tz)
// by particle id - 1, applied at every force evaluation from now on.
int grains_gpu_set_hydro_ft(grains_gpu_system* s, const double* ft, size_t m) {
  if (!s || !s->initialized || (!ft && m)) return GRAINS_ARGUMENT_ERROR;
  const int n = s->nparticles;
  if (m != (size_t)n)
    return fail(s, GRAINS_ARGUMENT_ERROR,
                "hydro force rows " + std::to_string(m) + " != particles " + std::to_string(n));
  HIPCHECK(s, hipSetDevice(s->device));
  if (n == 0) return GRAINS_OK;
  const bool fresh = !s->d_hyd;
  if (fresh) HIPCHECK(s, track(s, &s->d_hyd, 6 * (size_t)n));
  std::vector<int> ids((size_t)n);
  HIPCHECK(s, hipMemcpy(ids.data(), s->P.id, sizeof(int) * n, hipMemcpyDeviceToHost));
  std::vector<double> buf(6 * (size_t)n);
  for (int i = 0; i < n; ++i)
    for (int c = 0; c < 6; ++c) buf[(size_t)c * n + i] = ft[6 * (size_t)(ids[i] - 1) + c];
  HIPCHECK(s, hipStreamSynchronize(s->stream));
  HIPCHECK(s, hipMemcpy(s->d_hyd, buf.data(), sizeof(double) * buf.size(), hipMemcpyHostToDevice));
  // step graphs captured before hold StepParams without the fluid forces
  if (fresh) s->drop_graphs();
  return GRAINS_OK;
}

int grains_gpu_force_sums(grains_gpu_system* s, double* out) {
  if (!s || !s->initialized || !out) return GRAINS_ARGUMENT_ERROR;
  HIPCHECK(s, hipSetDevice(s->device));
  const int n = s->nparticles;
  if (!s->d_fsum) HIPCHECK(s, track(s, &s->d_fsum, (size_t)n));
  s->fsum_on = true;
  s->fsum_buf = s->d_fsum;
  int e = s->forces(kModeForcesOnly, 0, false);
  s->fsum_on = false;
  if (e) return e;
  if ((e = s->check_status())) return e;
  std::vector<int> ids((size_t)n);
  std::vector<double> buf((size_t)n);
  HIPCHECK(s, hipMemcpy(ids.data(), s->P.id, sizeof(int) * n, hipMemcpyDeviceToHost));
  HIPCHECK(s, hipMemcpy(buf.data(), s->d_fsum, sizeof(double) * n, hipMemcpyDeviceToHost));
  for (int i = 0; i < n; ++i) out[ids[i] - 1] = buf[i];
  return GRAINS_OK;
}

int grains_gpu_get_state(grains_gpu_system* s, double* x, double* v, double* w, double* wb, double* q, double* F,
                         double* T) {
  if (!s || !s->initialized) return GRAINS_ARGUMENT_ERROR;
  HIPCHECK(s, hipSetDevice(s->device));
  HIPCHECK(s, hipStreamSynchronize(s->stream));
  const int n = s->nparticles;
  std::vector<int> ids((size_t)n);
  HIPCHECK(s, hipMemcpy(ids.data(), s->P.id, sizeof(int) * n, hipMemcpyDeviceToHost));
  std::vector<double> buf((size_t)n);
  const int cx = s->cur & 1, cv = (s->cur >> 1) & 1;
  auto down = [&](double* dst, double* src, int stride, int comp) -> hipError_t {
    if (!dst) return hipSuccess;
    hipError_t e = hipMemcpy(buf.data(), src, sizeof(double) * n, hipMemcpyDeviceToHost);
    if (e != hipSuccess) return e;
    for (int i = 0; i < n; ++i) dst[(size_t)stride * (ids[i] - 1) + comp] = buf[i];
    return hipSuccess;
  };
  for (int a = 0; a < 3; ++a) {
    HIPCHECK(s, down(x, s->P.x[cx][a], 3, a));
    HIPCHECK(s, down(v, s->P.v[cv][a], 3, a));
    HIPCHECK(s, down(w, s->P.w[cv][a], 3, a));
    HIPCHECK(s, down(wb, s->P.wb[a], 3, a));
    HIPCHECK(s, down(F, s->P.F[a], 3, a));
    HIPCHECK(s, down(T, s->P.T[a], 3, a));
  }
  for (int a = 0; a < 4; ++a) HIPCHECK(s, down(q, s->P.q[a], 4, a));
  return s->check_status();
}

int grains_gpu_get_basis(grains_gpu_system* s, double* B) {
  if (!s || !s->initialized || !B) return GRAINS_ARGUMENT_ERROR;
  if (!s->has_shapes()) return fail(s, GRAINS_ARGUMENT_ERROR, "the system has no non-spherical particles");
  HIPCHECK(s, hipSetDevice(s->device));
  HIPCHECK(s, hipStreamSynchronize(s->stream));
  const int n = s->nparticles;
  std::vector<int> ids((size_t)n);
  HIPCHECK(s, hipMemcpy(ids.data(), s->P.id, sizeof(int) * n, hipMemcpyDeviceToHost));
  std::vector<double> buf((size_t)n);
  const int cx = s->cur & 1;
  for (int j = 0; j < 9; ++j) {
    HIPCHECK(s, hipMemcpy(buf.data(), s->P.B[cx][j], sizeof(double) * n, hipMemcpyDeviceToHost));
    for (int i = 0; i < n; ++i) B[9 * (size_t)(ids[i] - 1) + j] = buf[i];
  }
  return s->check_status();
}

int grains_gpu_set_basis(grains_gpu_system* s, const double* B) {
  if (!s || !s->initialized || !B) return GRAINS_ARGUMENT_ERROR;
  if (!s->has_shapes()) return fail(s, GRAINS_ARGUMENT_ERROR, "the system has no non-spherical particles");
  HIPCHECK(s, hipSetDevice(s->device));
  HIPCHECK(s, hipStreamSynchronize(s->stream));
  const int n = s->nparticles;
  std::vector<int> ids((size_t)n);
  HIPCHECK(s, hipMemcpy(ids.data(), s->P.id, sizeof(int) * n, hipMemcpyDeviceToHost));
  std::vector<double> buf((size_t)n);
  const int cx = s->cur & 1;
  for (int j = 0; j < 9; ++j) {
    for (int i = 0; i < n; ++i) buf[i] = B[9 * (size_t)(ids[i] - 1) + j];
    HIPCHECK(s, hipMemcpy(s->P.B[cx][j], buf.data(), sizeof(double) * n, hipMemcpyHostToDevice));
  }
  return GRAINS_OK;
}

int grains_gpu_set_state(grains_gpu_system* s, const double* x, const double* v, const double* w, const double* q,
                         int recompute_forces) {
  if (!s || !s->initialized) return GRAINS_ARGUMENT_ERROR;
  HIPCHECK(s, hipSetDevice(s->device));
  HIPCHECK(s, hipStreamSynchronize(s->stream));
  const int n = s->nparticles;
  std::vector<int> ids((size_t)n);
  HIPCHECK(s, hipMemcpy(ids.data(), s->P.id, sizeof(int) * n, hipMemcpyDeviceToHost));
  const int cx = s->cur & 1, cv = (s->cur >> 1) & 1;
  std::vector<double> buf((size_t)n);
  auto get = [&](double* src, std::vector<double>& dst, int stride, int comp) -> hipError_t {
    hipError_t e = hipMemcpy(buf.data(), src, sizeof(double) * n, hipMemcpyDeviceToHost);
    for (int i = 0; i < n; ++i) dst[(size_t)stride * (ids[i] - 1) + comp] = buf[i];
    return e;
  };
  auto put = [&](double* dst, const std::vector<double>& src, int stride, int comp) -> hipError_t {
    for (int i = 0; i < n; ++i) buf[i] = src[(size_t)stride * (ids[i] - 1) + comp];
    return hipMemcpy(dst, buf.data(), sizeof(double) * n, hipMemcpyHostToDevice);
  };
  // current q and w (needed to rebuild the body-fixed angular velocity)
  std::vector<double> qq((size_t)4 * n), ww((size_t)3 * n);
  for (int a = 0; a < 4; ++a) HIPCHECK(s, get(s->P.q[a], qq, 4, a));
  for (int a = 0; a < 3; ++a) HIPCHECK(s, get(s->P.w[cv][a], ww, 3, a));
  if (q) std::copy(q, q + 4 * (size_t)n, qq.begin());
  if (w) std::copy(w, w + 3 * (size_t)n, ww.begin());
  std::vector<double> wbv((size_t)3 * n);
  for (int i = 0; i < n; ++i) {
    quat qi{d3{qq[4 * i], qq[4 * i + 1], qq[4 * i + 2]}, qq[4 * i + 3]};
    d3 wb = to_body(qi, d3{ww[3 * i], ww[3 * i + 1], ww[3 * i + 2]});
    wbv[3 * i] = wb.x;
    wbv[3 * i + 1] = wb.y;
    wbv[3 * i + 2] = wb.z;
  }
  std::vector<double> tmp;
  for (int a = 0; a < 3; ++a) {
    if (x) {
      tmp.assign(x, x + 3 * (size_t)n);
      HIPCHECK(s, put(s->P.x[cx][a], tmp, 3, a));
    }
    if (v) {
      tmp.assign(v, v + 3 * (size_t)n);
      HIPCHECK(s, put(s->P.v[cv][a], tmp, 3, a));
    }
    HIPCHECK(s, put(s->P.w[cv][a], ww, 3, a));
    HIPCHECK(s, put(s->P.wb[a], wbv, 3, a));
  }
  for (int a = 0; a < 4; ++a) HIPCHECK(s, put(s->P.q[a], qq, 4, a));
  int e = s->bin_current();
  if (e) return e;
  if (recompute_forces) {
    if ((e = s->forces(kModeForcesOnly, 0, false))) return e;
  }
  return s->check_status();
}

int grains_gpu_construction_state(grains_gpu_system* s, double* x, double* v, double* w, double* basis) {
  if (!s) return GRAINS_ARGUMENT_ERROR;
  const size_t n = (size_t)s->nparticles;
  if (x) std::copy(s->h_x.begin(), s->h_x.begin() + 3 * n, x);
  if (v) std::copy(s->h_v.begin(), s->h_v.begin() + 3 * n, v);
  if (w) std::copy(s->h_w.begin(), s->h_w.begin() + 3 * n, w);
  if (basis) {
    if (s->h_B.size() < 9 * n) return fail(s, GRAINS_ARGUMENT_ERROR, "the system has no non-spherical particles");
    std::copy(s->h_B.begin(), s->h_B.begin() + 9 * n, basis);
  }
  return GRAINS_OK;
}

int grains_gpu_get_masses(grains_gpu_system* s, double* mass, double* inertia, double* weight) {
  if (!s) return GRAINS_ARGUMENT_ERROR;
  for (int i = 0; i < s->nparticles; ++i) {
    if (mass) mass[i] = s->h_mass[i];
    if (inertia) inertia[i] = s->h_inertia[i];
    if (weight)
      for (int a = 0; a < 3; ++a) weight[3 * i + a] = s->cfg.gravity[a] * s->h_wfac[i];
  }
  return GRAINS_OK;
}

int grains_gpu_set_neighbor_list(grains_gpu_system* s, int capacity, double skin) {
  if (!s || capacity < 0 || skin != skin) return GRAINS_ARGUMENT_ERROR;
  if (s->initialized) return fail(s, GRAINS_ARGUMENT_ERROR, "the neighbour list is sized before initialisation");
  s->nbr_cap = capacity;
  s->verlet_skin = skin;
  return GRAINS_OK;
}

int grains_gpu_neighbor_stats(grains_gpu_system* s, double* skin, long* builds) {
  if (!s || !s->initialized) return GRAINS_ARGUMENT_ERROR;
  int e = s->check_status();
  if (e) return e;
  if (skin) *skin = s->verlet ? s->V.skin : 0.;
  if (builds) *builds = s->verlet ? s->last_status.n_rebuild : 0;
  return GRAINS_OK;
}

int grains_gpu_set_contact_stats(grains_gpu_system* s, int on) {
  if (!s) return GRAINS_ARGUMENT_ERROR;
  s->stats_on = on != 0;
  if (s->initialized && on) {
    HIPCHECK(s, hipSetDevice(s->device));
    HIPCHECK(s, hipMemsetAsync(s->d_stat_tot, 0, 8 * sizeof(double), s->stream));
  }
  return GRAINS_OK;
}

int grains_gpu_contact_stats(grains_gpu_system* s, double* out) {
  if (!s || !s->initialized || !out) return GRAINS_ARGUMENT_ERROR;
  HIPCHECK(s, hipSetDevice(s->device));
  double t[8];
  HIPCHECK(s, hipMemcpyAsync(t, s->d_stat_tot, sizeof(t), hipMemcpyDeviceToHost, s->stream));
  int e = s->check_status();
  if (e) return e;
  out[0] = t[1];
  out[1] = t[3] > 0. ? t[0] / t[3] : 0.;
  out[2] = t[2];
  out[3] = t[5] > 0. ? t[4] / t[5] : 0.;
  out[4] = t[3];
  out[5] = t[5];
  return GRAINS_OK;
}

int grains_gpu_set_record_pairs(grains_gpu_system* s, int on) {
  if (!s) return GRAINS_ARGUMENT_ERROR;
  s->record_pairs = on ? 1 : 0;
  return GRAINS_OK;
}

int grains_gpu_get_pairs(grains_gpu_system* s, int* buf, size_t cap, size_t* n) {
  if (!s || !s->initialized || !n) return GRAINS_ARGUMENT_ERROR;
  int e = s->check_status();
  if (e) return e;
  size_t np = (size_t)std::min(s->last_status.npairs, s->pair_cap);
  *n = np;
  if (buf && cap) {
    size_t m = std::min(np, cap);
    HIPCHECK(s, hipMemcpy(buf, s->d_pairs, sizeof(int) * 2 * m, hipMemcpyDeviceToHost));
  }
  return GRAINS_OK;
}

int grains_gpu_contact_counts(grains_gpu_system* s, size_t* contacts, size_t* sliding) {
  if (!s || !s->initialized) return GRAINS_ARGUMENT_ERROR;
  int e = s->check_status();
  if (e) return e;
  if (contacts) *contacts = (size_t)s->last_status.npairs;
  if (sliding) *sliding = (size_t)s->last_status.n_slip;
  return GRAINS_OK;
}

int grains_gpu_time_interval(const grains_gpu_system* s, double* t_start, double* t_end, double* dt) {
  if (!s) return GRAINS_ARGUMENT_ERROR;
  if (t_start) *t_start = s->cfg.t_start;
  if (t_end) *t_end = s->t_end;
  if (dt) *dt = s->cfg.dt;
  return GRAINS_OK;
}

size_t grains_gpu_num_particles(const grains_gpu_system* s) { return s ? (size_t)s->nparticles : 0; }
double grains_gpu_time(const grains_gpu_system* s) { return s ? s->time : 0.; }
void grains_gpu_set_time(grains_gpu_system* s, double t) {
  if (s) s->time = t;
}

int grains_gpu_grid(const grains_gpu_system* s, int* dims, double* edge) {
  if (!s || !s->initialized) return GRAINS_ARGUMENT_ERROR;
  for (int a = 0; a < 3; ++a) {
    if (dims) dims[a] = s->grid.n[a];
    if (edge) edge[a] = s->grid.edge[a];
  }
  return GRAINS_OK;
}

int grains_gpu_error_ids(grains_gpu_system* s, int* id0, int* id1) {
  if (!s) return GRAINS_ARGUMENT_ERROR;
  if (id0) *id0 = s->last_status.id0;
  if (id1) *id1 = s->last_status.id1;
  return GRAINS_OK;
}

int grains_gpu_recoveries(grains_gpu_system* s, long* n) {
  if (!s || !n) return GRAINS_ARGUMENT_ERROR;
  *n = s->last_status.n_recovered;
  return GRAINS_OK;
}

const char* grains_gpu_last_error(const grains_gpu_system* s) { return s ? s->last_error.c_str() : "null system"; }

}  // extern "C"
