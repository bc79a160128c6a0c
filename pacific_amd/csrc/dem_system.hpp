// The system object behind the opaque grains_gpu_system handle.
#pragma once
#include <hip/hip_runtime.h>

#include <functional>
#include <memory>
#include <string>
#include <vector>

#include "../../include/grains_gpu.h"
#include "dem_types.hpp"

namespace dem {
struct AdjointData;
}
namespace grains_insert {
struct Plan;
}

struct grains_gpu_system {
  grains_gpu_config cfg{};
  int device = 0;
  int nparticles = 0;
  double time = 0.;
  double t_end = 0.;  // <TimeInterval End> of an XML-built system
  // outputs of the time loop: <TimeSave> times, <RestartFile Name> (root of
  // _VelocityMaxMean.dat), <RawData Name WritingMode>
  std::vector<double> save_times;
  std::string save_root, rawdata_root;
  bool rawdata_binary = false;
  // particle class of each particle (index of its <Particle> in the input;
  // the RawData particle type), id order, and the class of the next ones
  std::vector<int> h_class;
  int next_class = 0;
  long steps_done = 0;
  bool initialized = false;
  int cur = 0;  // bit 0: position buffer, bit 1: velocity buffer
  int record_pairs = 0;
  bool needs_rebin = false;  // a launch sequence was interrupted
  int pair_cap = 0;
  std::string last_error;
  dem::DevStatus last_status{};

  // host copies of the construction data, id order
  std::vector<double> h_rad, h_mass, h_inertia, h_crust, h_wfac, h_cfac, h_req;
  std::vector<double> h_x, h_v, h_w, h_wb, h_q;
  std::vector<int> h_mat;
  // non-spherical particles: shape classes, their vertex table, and per
  // particle the class (-1 = sphere added by add_spheres), basis rows and
  // principal moments of inertia
  std::vector<dem::ShapeDesc> shapes;
  std::vector<double> shape_volume, shape_inertia;  // per unit density (1 / 6 per class)
  std::vector<double> shape_verts;
  std::vector<int> h_shape;
  std::vector<double> h_B, h_I3;
  std::vector<double> h_density;
  std::vector<std::vector<int>> shape_faces;  // per class: face vertex lists (polyhedra)
  // the Transform basis is also tracked for spheres (coupled runs report it)
  bool track_basis = false;
  bool has_shapes() const { return !shapes.empty() || track_basis; }
  // coupling with a fluid solver (GrainsCoupledWithFluid): the acceleration
  // is corrected by (1 - rho_f / rho) unless switched off, and the DEM time
  // step follows <TimeStep mindt maxdt ndt>
  bool fluid_corrected = true;
  double coupled_min_dt = 0., coupled_max_dt = 0.;
  size_t coupled_ndt = 0;
  std::vector<int> obs_to_fluid;
  // names the reload files carry (set by the XML front-end; defaults
  // otherwise): materials by index, polyhedron files by shape class,
  // obstacles by index; saves made so far (A/B alternation)
  std::vector<std::string> mat_names, shape_files, obs_names;
  int reload_count = 0;
  std::vector<dem::BoxObstacle> obstacles;
  dem::HookeModel models[dem::kMaxMaterials * dem::kMaxMaterials]{};
  dem::HookeModel default_model{};
  bool has_default = false;

  // device
  dem::Grid grid{};
  dem::Particles P{};
  dem::Sorted S{};
  dem::BoxObstacle* d_obs = nullptr;
  dem::ShapeDesc* d_shapes = nullptr;
  double* d_verts = nullptr;
  dem::HookeModel* d_models = nullptr;
  dem::DevStatus* d_status = nullptr;
  int* d_pairs = nullptr;
  // per-step neighbour candidate list (0 = the force sweep searches the 27
  // cells itself); set before initialisation
  int nbr_cap = 32;
  dem::GjkOpts gjk_opts{0, 0, 0, 0, 1e-10};  // <CollisionDetectionAlgorithm>
  int* d_nbr_count = nullptr;
  unsigned* d_nbr = nullptr;
  dem::NbrList nbr_list() const { return dem::NbrList{d_nbr_count, d_nbr, nbr_cap, nparticles}; }
  // Verlet list: candidates within the contact distance plus a skin, binned
  // on a grid of its own and rebuilt only once some particle has moved more
  // than half the skin (requested skin < 0: a tenth of the linked-cell size;
  // 0: the candidate list is rebuilt every step)
  double verlet_skin = -1.;
  bool verlet = false;  // decided at initialisation
  dem::VerletState V{};
  // contacts of the current step filtered from the Verlet list
  static constexpr int kContactCap = 16;
  int* d_ct_count = nullptr;
  unsigned* d_ct = nullptr;
  dem::NbrList contact_list() const {
    return verlet ? dem::NbrList{d_ct_count, d_ct, kContactCap, nparticles} : dem::NbrList{};
  }

  // contacts found by the narrow phase of a system with non-spherical
  // bodies, read by its force sweep (allocated when the system has any)
  static constexpr int kFeatCap = 32;
  dem::ContactFeat cf{};
  // per-particle sums of the contact-force magnitudes of one evaluation
  // (grains_gpu_force_sums; allocated on first use)
  // forces and torques of the fluid on each particle (AppPRSHydroFT), 6 x n
  // by storage index; allocated by the first grains_gpu_set_hydro_ft
  double* d_hyd = nullptr;
  double* d_fsum = nullptr;
  bool fsum_on = false;
  double* fsum_buf = nullptr;  // buffer of the evaluation when fsum_on (d_fsum or a caller's scratch)

  // contact statistics (AppCollision): per-block partials of the last force
  // launch and running totals; force_time = time of the next force launch
  // HookeMemory contact memory (allocated when a model has memory)
  static constexpr int kMemoryCap = 16;
  dem::ContactMemory mem{};
  bool has_memory() const;
  bool stats_on = false;
  double* d_stat_part = nullptr;
  double* d_stat_tot = nullptr;
  double force_time = 0.;
  // stale flag of the Verlet list: set by the drift that finds the list
  // stale, read by the rebuild launches, cleared by the next contact filter
  // (the same kernel arguments every step, so that steps can be replayed
  // from a captured graph)
  static constexpr int kStale = 1;
  int next_stamp() const { return kStale; }
  dem::VerletState vstate(int st) const {
    dem::VerletState v = V;
    v.stamp = verlet ? st : 0;
    return v;
  }
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  std::vector<void*> allocations;

  ~grains_gpu_system();
  int initialize();
  int build_grid();
  int allocate();
  int upload();
  void free_device();
  dem::StepParams params() const;
  int bin_current();
  int reset_pair_counter();
  // commit: the evaluation belongs to the time loop (contact memory advances)
  int forces(int mode, int first_step, bool commit = true);
  // optional per-launch HIP events (force sweeps and binning sequences)
  struct Profile {
    std::vector<hipEvent_t> force_ev, bin_ev, detect_ev, narrow_ev;
    size_t nforce = 0, nbin = 0, ndetect = 0, nnarrow = 0;
  };
  Profile* prof = nullptr;
  int mark(std::vector<hipEvent_t>& evs, size_t& n);
  int bin_launch();
  // adjoint support: checkpoints of a forward run, and a recorder that keeps
  // binned (x_{n+1}, v_{n+1/2}, w_{n+1/2}) of every LeapFrog step
  dem::AdjointData* adj = nullptr;
  bool recording = false;
  int rec_slot = 0;
  int record_state();
  void free_adjoint();
  int enqueue_steps(int nsteps);
  int enqueue_steps_impl(int nsteps);
  // LeapFrog steps replayed from captured graphs, kGraphSteps fused steps
  // per graph (even, so that the buffer parities return), one graph per
  // parity state at capture (cur, mem.cur), all dropped when dt changes
  static constexpr int kGraphSteps = 64;
  std::vector<std::pair<int, hipGraphExec_t>> graphs;
  double graph_dt = 0.;
  bool graphs_failed = false;  // capture refused on this stack: plain launches
  std::string graph_error;     // why
  long graph_replays = 0;
  bool graph_steps_allowed() const;
  int graph_chunk(bool& done);
  void drop_graphs();
  int recover();
  int check_status();
  // particles of the XML input waiting for insertion over time
  // (<Mode Type="OverTime">), tried once per step by the time loop
  std::shared_ptr<grains_insert::Plan> insertion;
  // Adds one particle to an initialised system: the state goes back to the
  // construction data, `add` appends the particle, the device state is built
  // again from them, and the forces and torques carried to the next step's
  // first half kick are put back
  int grow(const std::function<int()>& add);
};
