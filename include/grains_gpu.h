/*
 * grains_gpu.h -- C ABI of the MI355X DEM time step (libgrains_gpu.so).
 *
 * Plain C types only: pointers, sizes, doubles and ints.  All arrays are
 * host memory, row-major, ordered by particle id (ids are 1..N in insertion
 * order, obstacles are -1..-M), so that a caller written against Grains3D's
 * Particle / Obstacle / ContactForceModel objects can fill them directly.
 *
 * Every call returns GRAINS_OK (0) or an error code; grains_gpu_last_error()
 * gives the message.  Errors raised inside the time step (ContactError,
 * MotionError, SimulationError of the reference, GRAINS_V1/Grains/Base/
 * include/Error.hh) are flagged on the device and returned by the next call
 * that synchronises (grains_gpu_step, grains_gpu_synchronize, the getters).
 *
 * A system is not thread safe; distinct systems are independent and may run
 * on distinct threads.
 */
#ifndef GRAINS_GPU_H
#define GRAINS_GPU_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct grains_gpu_system grains_gpu_system;

/* status codes; 1-3 mirror the reference's ContactError / MotionError /
 * SimulationError (Grains.cpp:423-450) */
enum {
  GRAINS_OK = 0,
  GRAINS_CONTACT_ERROR = 1,
  GRAINS_MOTION_ERROR = 2,
  GRAINS_SIMULATION_ERROR = 3,
  GRAINS_CONFIG_ERROR = 4,
  GRAINS_DEVICE_ERROR = 5,
  GRAINS_ARGUMENT_ERROR = 6
};

/* arithmetic of the force model */
enum { GRAINS_FP64 = 0, GRAINS_MIXED = 1 };

/* time integrators, <TimeIntegration Type=...> (TimeIntegratorBuilderFactory.cpp:13-40) */
enum {
  GRAINS_LEAPFROG = 0,               /* "SecondOrderLeapFrog" (default) */
  GRAINS_FIRST_ORDER_EXPLICIT = 1,   /* "FirstOrderExplicit" */
  GRAINS_SECOND_ORDER_EXPLICIT = 2,  /* "SecondOrderExplicit" */
  GRAINS_ADAMS_BASHFORTH2 = 3        /* "SecondOrderAdamsBashforth" */
};

typedef struct {
  double origin[3];        /* <Origin OX OY OZ> */
  double max[3];           /* <LinkedCell MX MY MZ> (maximum coordinates) */
  int periodic[3];         /* <Periodicity PX PY PZ> */
  double cell_size_factor; /* <LinkedCell CellSizeFactor> (>= 1) */
  double gravity[3];       /* <Forces><Gravity GX GY GZ> */
  double fluid_density;    /* 0 for a dry simulation */
  double dt;               /* <TimeStep dt> */
  double t_start;          /* <TimeInterval Start> */
  int scheme;              /* GRAINS_LEAPFROG ... */
  int precision;           /* GRAINS_FP64 or GRAINS_MIXED */
  int device;              /* HIP device ordinal */
} grains_gpu_config;

/* Replaces Grains::Construction / Forces / AdditionalFeatures setup of an
 * empty system (GRAINS_V1/Grains/Base/src/Grains.cpp:535-1100). */
int grains_gpu_create(const grains_gpu_config* cfg, grains_gpu_system** out);

/* Builds a system from a Grains3D XML input file (the <Construction>,
 * <Forces> and <Simulation> sections, Grains.cpp:95-179).  Particles listed
 * through <ParticlePosition><File> or <StructuredArray> are inserted at the
 * initial time; the system is initialised (grains_gpu_initialize) on return.
 * Replaces GrainsBuilderFactory::create + Grains::do_before_time_stepping. */
int grains_gpu_create_from_xml(const char* xml_path, int precision, grains_gpu_system** out);
/* The same input read up to initialisation: the construction data (the
 * particles inserted at the initial time) without the device, for
 * inspection with grains_gpu_construction_state.  No counterpart in the
 * reference, whose Grains::Construction always goes on to the run; a
 * <Reload> is not read. */
int grains_gpu_parse_xml(const char* xml_path, int precision, grains_gpu_system** out);
/* Construction data of a system (id order, before or after initialisation:
 * the values it was built from): x[3n], v[3n], w[3n], basis[9n] (NULL to
 * skip; basis only for systems with non-spherical particles). */
int grains_gpu_construction_state(grains_gpu_system* sys, double* x, double* v, double* w, double* basis);
/* Insertion over time (<ParticleInsertion><Mode Type="OverTime">,
 * Grains.cpp:322-327, insertParticle :2070-2140): one insertion call of a
 * time step, made before the step; the new particle (if one went in) gets
 * the next id.  grains_gpu_run makes these calls itself.  The number of
 * particles still waiting: */
int grains_gpu_try_insert(grains_gpu_system* sys, int* inserted);
int grains_gpu_pending_insertions(grains_gpu_system* sys, long* n);

void grains_gpu_destroy(grains_gpu_system* sys);

/* Hooke contact model of a material pair; matB < 0 sets the Default model.
 * Parameters of <Hooke><kn/><en/><muc/><kr/>[<etat/>]</Hooke>
 * (HookeContactForceModel.cpp:170-199); etat = -1 selects the default
 * tangential damping.  Replaces ContactBuilderFactory::define
 * (ContactBuilderFactory.cpp:82-117). */
int grains_gpu_set_contact_model(grains_gpu_system* sys, int matA, int matB, double kn, double en, double etat,
                                 double muc, double kr);

/* HookeMemory contact model for a material pair (HookeMemoryContactForceModel,
 * HookeMemoryContactForceModel.cpp:11-29, parameters :253-307):
 * p = {kn, en, etat (-1: damping as the normal one), muc, kt, mur (0: no
 * rolling friction), etarpf, epst}.  Each particle keeps, per partner met
 * at the last evaluation of the time loop, the cumulative tangential spring,
 * the previous normal and the rolling spring torque (at most 16 partners; a
 * particle with more stops the run with GRAINS_DEVICE_ERROR).  Before
 * grains_gpu_initialize. */
int grains_gpu_set_memory_contact_model(grains_gpu_system* sys, int matA, int matB, const double* p);

/* Adds n spheres in insertion order (ids continue from the current count).
 * radius, density, crust: n values each; material: n ints; x, v, w: 3n
 * values (w = space-fixed angular velocity); v and w may be NULL (zero).
 * Replaces Particle(DOMNode*) + Grains::insertParticle
 * (Particle.cpp:150-166, Grains.cpp:2070-2140). */
int grains_gpu_add_spheres(grains_gpu_system* sys, size_t n, const double* radius, const double* density,
                           const double* crust, const int* material, const double* x, const double* v,
                           const double* w);

/* Non-spherical particles.  A shape class is a convex with its crust
 * thickness, <Convex CrustThickness><Sphere Radius/> | <Box LX LY LZ/> |
 * <Polyhedron Name/></Convex> (ConvexBuilderFactory, RigidBodyWithCrust.cpp:
 * 132-168; Sphere.cpp:61-81, Box.cpp:181-279, Polyhedron.cpp:169-359):
 * type 0 sphere (radius), 1 box (edge lengths L[3]), 2 polyhedron (nv
 * vertices, 3 doubles each, and nf faces given by face_sizes[nf] and the
 * concatenated vertex indices).  *shape_id receives the class index. */
int grains_gpu_add_shape(grains_gpu_system* sys, int type, double radius, const double* L, int nv,
                         const double* verts, int nf, const int* face_sizes, const int* face_idx, double crust,
                         int* shape_id);
/* Adds n particles of a shape class (ids continue from the current count):
 * density and material per particle, x, v, w as for spheres, basis = 9
 * doubles per particle (row-major rotation of <AngularPosition>, NULL =
 * identity) from which the quaternion follows (Quaternion.cpp:200-255).
 * Contacts of non-spherical bodies go through GJK on the crust-shrunk
 * shapes (RigidBodyWithCrust.cpp:277-347); their rotation follows
 * ParticleKinematics3D (ParticleKinematics3D.cpp:56-77).  SecondOrderLeapFrog
 * only.  Replaces Particle(DOMNode*) for non-spherical convexes. */
int grains_gpu_add_particles(grains_gpu_system* sys, int shape, size_t n, const double* density,
                             const int* material, const double* x, const double* v, const double* w,
                             const double* basis);
/* out = {volume, circumscribed radius, inertia xx xy xz yy yz zz, crust
 * scaling[3], bounding-box half extents[3]}, per unit density */
int grains_gpu_shape_info(grains_gpu_system* sys, int shape, double* out);
/* <CollisionDetectionAlgorithm> of Grains::Construction (Grains.cpp:640-770;
 * GrainsExec::m_colDetGJK_SV, m_colDetTolerance, m_colDetAcceleration,
 * m_colDetBoundingVolume), before initialisation: gjk_sv 1 = the
 * signed-volume sub-distance (GJK_SV.cpp:428-525) for the general pairs,
 * 0 = Johnson's (the ContactError recovery uses Johnson's either way);
 * tolerance = relative tolerance of the optimality gap (>= 1e-10, default
 * 1e-10); acceleration 1 = Nesterov; bounding_volume 1 = OBB, 2 = OBC
 * pre-test of general pairs (OBB.cpp:144, OBC.cpp:226), 0 = none. */
int grains_gpu_set_collision(grains_gpu_system* sys, int gjk_sv, double tolerance, int acceleration,
                             int bounding_volume);
/* The device narrow phase evaluated on the host for two bodies of shape
 * classes shapeA, shapeB at centres x and row-major bases B
 * (RigidBodyWithCrust::ClosestPoint and its ContactError handling,
 * RigidBodyWithCrust.cpp:205-354, 534-590): out = {contact point[3],
 * overlap vector[3], overlap distance, status} with status 1 features set,
 * 0 no contact, -1 ContactError, 2 recovered.  For checks without a device. */
int grains_gpu_closest_point(grains_gpu_system* sys, int shapeA, const double* xA, const double* BA, int shapeB,
                             const double* xB, const double* BB, double* out);
/* Transform bases of all particles (id order, 9 per particle, row-major) */
int grains_gpu_get_basis(grains_gpu_system* sys, double* basis);
int grains_gpu_set_basis(grains_gpu_system* sys, const double* basis);

/* Adds a static box obstacle: edge lengths L[3], centre[3], basis[9] (row
 * major rotation, NULL = identity).  Replaces SimpleObstacle(DOMNode*) with
 * a <Box LX LY LZ> convex (SimpleObstacle.cpp:40-90, Box.cpp:52-60). */
int grains_gpu_add_box(grains_gpu_system* sys, const double* L, const double* centre, const double* basis,
                       double crust, int material);

/* Builds the linked-cell grid, uploads the state, bins it and (LeapFrog)
 * computes the initial forces (Grains.cpp:95-179, 2522-2527). */
int grains_gpu_initialize(grains_gpu_system* sys);

/* Runs nsteps time steps of Grains::Simulation's loop body
 * (Grains.cpp:286-452) without output. */
int grains_gpu_step(grains_gpu_system* sys, int nsteps);

/* Same, timed with HIP events on the system's stream; *ms = elapsed time. */
int grains_gpu_step_timed(grains_gpu_system* sys, int nsteps, double* ms);

/* Same, with HIP events around every force sweep and every binning sequence
 * on the system's stream: out[0] total ms, out[1] summed force-sweep ms,
 * out[2] number of force sweeps, out[3] summed binning ms, out[4] number of
 * binning sequences, out[5] summed neighbour-search ms, out[6] number of
 * neighbour searches, out[7] summed narrow-phase ms of non-spherical bodies,
 * out[8] number of those narrow phases (out holds 9 doubles).  No counterpart in the reference (instrumentation of
 * the Grains::Simulation timers, solvercomputingtime.cpp:48-215). */
int grains_gpu_profile_steps(grains_gpu_system* sys, int nsteps, double* out);

int grains_gpu_synchronize(grains_gpu_system* sys);

/* Step batches replay their fused LeapFrog steps from captured HIP graphs
 * (64 steps each, captured on first use per buffer-parity state).  Reports
 * the graphs held (-1: capture refused on this stack, plain launches; why
 * receives the reason) and the replays so far.  No counterpart in the
 * reference (launch mechanics of this library). */
int grains_gpu_graph_stats(grains_gpu_system* sys, int* ngraphs, long* replays, char* why, size_t cap);

/* Forces and torques at the current state (Grains::computeParticlesForceAndTorque,
 * Grains.cpp:459-482); results via grains_gpu_get_state(F, T). */
int grains_gpu_compute_forces(grains_gpu_system* sys);
/* Forces at the current state, as grains_gpu_compute_forces, and for each
 * particle (id order, n doubles) the sum of the magnitudes of the contact
 * forces it received: the normaliser of per-particle force-error figures.
 * No counterpart in the reference (a diagnostic of this library). */
int grains_gpu_force_sums(grains_gpu_system* sys, double* out);
/* The same sums into the caller's DEVICE buffer d_out (n doubles, id order),
 * allocating nothing: no counterpart in the reference; for callers that
 * cannot allocate device memory (code under HIP graph capture, serving
 * loops, harnesses that forbid hipMalloc).  The per-particle partial sums
 * live in a static device array of 4,194,304 doubles (32 MiB per GPU, held
 * until the process ends): larger systems are refused with
 * GRAINS_ARGUMENT_ERROR.  Every call on one GPU shares that array, so calls
 * that overlap on one GPU (several threads, several systems on different
 * streams) race; calls in sequence are safe. */
int grains_gpu_force_sums_noalloc(grains_gpu_system* sys, double* d_out);

/* State in id order: x, v, w, wb, F, T: 3 doubles per particle; q: 4 (x, y,
 * z, w).  NULL pointers are skipped. */
int grains_gpu_get_state(grains_gpu_system* sys, double* x, double* v, double* w, double* wb, double* q,
                         double* F, double* T);

/* Overwrites positions, velocities, angular velocities and quaternions (id
 * order; NULL keeps the current value), re-bins, and recomputes the forces
 * when recompute_forces != 0 (the LeapFrog state after a reload). */
int grains_gpu_set_state(grains_gpu_system* sys, const double* x, const double* v, const double* w,
                         const double* q, int recompute_forces);

/* mass, sphere moment of inertia and weight (3 per particle) as built from
 * radius and density (Particle.cpp:157-166, 413-418). */
int grains_gpu_get_masses(grains_gpu_system* sys, double* mass, double* inertia, double* weight);

/* Neighbour candidates of the broad phase of LinkedCell::ComputeForces
 * (LinkedCell.cpp:1019-1103), before grains_gpu_initialize only.
 * capacity: listed candidates per particle (default 32; a particle with more
 * is searched in its cells instead); 0 = no list, the force sweep searches
 * the 27 cells itself.  skin = 0: the list is rebuilt every step, after the
 * linked-cell binning, and forces are bit-identical to the cell search.
 * skin > 0 (default, skin < 0: a tenth of the cell size 2 Rmax factor;
 * spheres only): a Verlet list of every pair closer than the contact
 * distance plus the skin, binned on its own grid and rebuilt only once some
 * particle has moved more than half the skin, so that the contacts of every
 * step are among the listed pairs: the same contacts and per-contact
 * arithmetic, contact sums in the list's order. */
int grains_gpu_set_neighbor_list(grains_gpu_system* sys, int capacity, double skin);
/* Skin of the Verlet list in use (0: none) and the number of list builds
 * so far.  No counterpart in the reference. */
int grains_gpu_neighbor_stats(grains_gpu_system* sys, double* skin, long* builds);

/* Contact statistics of the time loop (AppCollision::addToContactsFeatures,
 * AppCollision.cpp:159-189, reported by Grains::Simulation,
 * Grains.cpp:228-240): off by default; turning them on clears them.  A
 * particle pair counts as often per force evaluation as the reference meets
 * it: once, or once per pair of its periodic images (master and clones in
 * the ghost layers) that exist together; a particle-obstacle pair once
 * (the reference's clone-obstacle meetings are not counted).  out[0] maximum overlap,
 * out[1] mean overlap, out[2] time of the maximum (first reached), out[3]
 * mean GJK iterations (closed-form contact points count 1), out[4] overlaps
 * counted, out[5] closest-point evaluations with iterations.  The means are
 * sums over counts (the reference updates a running mean). */
int grains_gpu_set_contact_stats(grains_gpu_system* sys, int on);
int grains_gpu_contact_stats(grains_gpu_system* sys, double* out);

/* Contact records of the last force evaluation: pairs (receiver id, partner
 * id), one per particle that received a contact force from the partner
 * (obstacle partners have negative ids).  Recording is off by default. */
int grains_gpu_set_record_pairs(grains_gpu_system* sys, int on);
int grains_gpu_get_pairs(grains_gpu_system* sys, int* buf, size_t cap, size_t* n);

/* ---- Output of the time loop (host side, over the entry points above) ---- */

/* Particle class of the particles added next (default 0): the index of
 * their <Particle> entry, written as the RawData particle type
 * (Particle::getGeometricType, Particle.cpp:1651-1654; set by
 * AllComponents.cpp:955).  Before grains_gpu_initialize. */
int grains_gpu_set_particle_class(grains_gpu_system* sys, int cls);

/* Maximum and mean translational speed of the particles
 * (AllComponents::ComputeMaxMeanVelocity, AllComponents.cpp:1781-1821;
 * obstacles do not move here). */
int grains_gpu_velocity_max_mean(grains_gpu_system* sys, double* vmax, double* vmean);

/* Coordination numbers in id order: the number of distinct components
 * (particles and obstacles) each particle was in contact with at the last
 * force evaluation (Particle::addContactingComponentID, Particle.cpp:
 * 1702-1706, reset by InitializeForce :1408-1415).  Needs contact records
 * (grains_gpu_set_record_pairs) during that evaluation. */
int grains_gpu_coordination_numbers(grains_gpu_system* sys, int* out);
/* The same at the current state, without records of the last step: one
 * force evaluation with contact records, after which the torsor the time
 * loop holds is restored (so that the next step is unchanged).  Serves the
 * post-processing calls of a coupled solver (SaveResults_Grains). */
int grains_gpu_coordination_at_state(grains_gpu_system* sys, int* out);
/* The post-processing writers of the input file (<PostProcessing><Writers>
 * <RawData>) at the current state, coordination numbers included: start != 0
 * starts the files (AllComponents::PostProcessing_start), otherwise a frame
 * is appended (AllComponents::PostProcessing); GrainsCoupledWithFluid::
 * InitialPostProcessing / doPostProcessing, GrainsCoupledWithFluid.cpp:
 * 870-893.  Nothing to do without a writer. */
int grains_gpu_post_processing(grains_gpu_system* sys, int start);

/* Reload (restart) files, __Format2024__ layout (Grains::saveReload,
 * Grains.cpp:2543-2584; AllComponents::write, AllComponents.cpp:1299-1415;
 * Particle::write2014 / write2014_binary, Particle.cpp:1167-1193;
 * ContactBuilderFactory::save, ContactBuilderFactory.cpp:270-329): writes
 * <root>A or <root>B (alternating per system, A first) with the particle
 * types, obstacles and file table; the particle records <root>X_Particles
 * (text, %.16e) or <root>X_Particles.bin (hybrid != 0: the binary record
 * of write2014_binary, preceded by its byte count); <root>_MatContact.xml at
 * the first save; one line "time <root>X" appended to <root>_RFTable.txt. */
int grains_gpu_save_reload(grains_gpu_system* sys, const char* root, int hybrid);
/* Reads a reload file written for the same particles (ids 1..n): time,
 * positions, bases, velocities, quaternions, the torsor the LeapFrog
 * records carry (no new force evaluation, as Grains.cpp:169-174 skips it
 * on reload), Adams-Bashforth data and contact memory (Grains.cpp:784-844,
 * AllComponents::read / read_particles, AllComponents.cpp:1024-1290). */
int grains_gpu_load_reload(grains_gpu_system* sys, const char* file);
/* The same from the last file listed in <RestartFile Name>_RFTable.txt of
 * the system's input file (<Reload Type="same">; GrainsCoupledWithFluid::
 * setReloadSame; GrainsExec::restartFileName_AorB, GrainsExec.cpp:455-469). */
int grains_gpu_reload_same(grains_gpu_system* sys);

/* One RawData frame at the current time (RawDataPostProcessingWriter::
 * one_output_Standard, RawDataPostProcessingWriter.cpp:362-600): the files
 * <root>_position_{x,y,z}.dat, _translational_velocity_{x,y,z}.dat,
 * _angular_velocity_{x,y,z}.dat, _coordinationNumber.dat and
 * _particleType.dat (last frame only); text (scientific, 6 decimals) or
 * binary.  truncate != 0 starts the files (ReloadType "new"), else the
 * frame is appended.  coord: coordination numbers in id order, or NULL
 * for zeros. */
int grains_gpu_write_rawdata(grains_gpu_system* sys, const char* root, int binary, int truncate, const int* coord);

/* Output settings of grains_gpu_run (the XML front-end sets them from
 * <TimeSave Start End Every>, <RestartFile Name> and <RawData Name
 * WritingMode>): output times, root of <root>_VelocityMaxMean.dat (NULL or
 * "": none), RawData root (NULL or "": none). */
int grains_gpu_set_outputs(grains_gpu_system* sys, const double* save_times, size_t nsave, const char* save_root,
                           const char* rawdata_root, int rawdata_binary);

/* Grains::Simulation to <TimeInterval End> with its outputs (Grains.cpp:
 * 279-452 and the initial and final outputs :130-146, :199-211): at the
 * start, at every output time and at the end a line "t  max|v|  mean|v|" in
 * <save_root>_VelocityMaxMean.dat; at the start and at every output time a
 * RawData frame.  Restart files are not written. */
int grains_gpu_run(grains_gpu_system* sys, int verbose);

size_t grains_gpu_num_particles(const grains_gpu_system* sys);
/* <TimeInterval Start End> and <TimeStep dt> of the system */
int grains_gpu_time_interval(const grains_gpu_system* sys, double* t_start, double* t_end, double* dt);
double grains_gpu_time(const grains_gpu_system* sys);
void grains_gpu_set_time(grains_gpu_system* sys, double t);
/* Directed contact records and, of those, contacts on the Coulomb cap
 * (|FT| = mu |FN|) of the last force evaluation made with recording on
 * (AppCollision::addToContactsFeatures keeps running contact features,
 * AppCollision.cpp:159-189). */
int grains_gpu_contact_counts(grains_gpu_system* sys, size_t* contacts, size_t* sliding);
/* grid cells per axis (GPU interior cells) */
int grains_gpu_grid(const grains_gpu_system* sys, int* dims, double* edge);
/* ids of the bodies named by the last device-side error */
int grains_gpu_error_ids(grains_gpu_system* sys, int* id0, int* id1);
/* number of ContactError recoveries (RigidBodyWithCrust.cpp:534-590) */
int grains_gpu_recoveries(grains_gpu_system* sys, long* n);
const char* grains_gpu_last_error(const grains_gpu_system* sys);

/* ---- Batched ensembles ---------------------------------------------------
 * Many independent systems (each with its own box, obstacles, contact
 * parameters and size) advanced by one persistent launch, one workgroup per
 * system.  No counterpart in the reference, which runs one system per process
 * (static state in GrainsExec, ContactBuilderFactory and Cell); each member
 * evolves exactly as grains_gpu_step would evolve it alone. */
typedef struct grains_gpu_batch grains_gpu_batch;

/* Members are configured with grains_gpu_create / set_contact_model /
 * add_spheres / add_box and NOT initialised; the batch copies their data (the
 * members may be destroyed afterwards).  SecondOrderLeapFrog only.  Bins every
 * member and computes its initial forces. */
int grains_gpu_batch_create(grains_gpu_system* const* systems, int nsys, int precision, grains_gpu_batch** out);
void grains_gpu_batch_destroy(grains_gpu_batch* b);
/* launch shape: 0 = flat mode (one launch per phase over the particles of
 * every member), or the workgroup size per member of the persistent launch,
 * a multiple of 64 from 64 to 512 */
int grains_gpu_batch_set_block(grains_gpu_batch* b, int block);
/* nsteps steps of every member; whole chunks of `chunk` steps replay one
 * captured hipGraph of the step's launches (chunk <= 0: direct launches) */
int grains_gpu_batch_step(grains_gpu_batch* b, int nsteps, int chunk);
int grains_gpu_batch_step_timed(grains_gpu_batch* b, int nsteps, int chunk, double* ms);
/* times trial_steps steps in each launch shape (0 = flat, then workgroups
 * of 64, 128, 192, ..., 512 -> ms_out[0..8]), keeps the fastest; advances
 * the members by 9 * trial_steps */
int grains_gpu_batch_autotune(grains_gpu_batch* b, int trial_steps, int* best_block, double* ms_out);
/* particles of member isys, or of the whole batch for isys < 0 */
size_t grains_gpu_batch_num_particles(const grains_gpu_batch* b, int isys);
double grains_gpu_batch_time(const grains_gpu_batch* b);
/* state of member isys in its own id order (NULL pointers skipped) */
int grains_gpu_batch_get_state(grains_gpu_batch* b, int isys, double* x, double* v, double* w, double* q, double* F,
                               double* T);
const char* grains_gpu_batch_last_error(const grains_gpu_batch* b);

/* ---- Coupling with a fluid solver ----------------------------------------
 * Handle-based form of the Basilisk interface (InterfaceGrains3DBasilisk.h),
 * returning status codes instead of ending the process. */

/* Builds a system from a coupled input file: <TimeStep mindt maxdt ndt>,
 * fluid density fluid_rho, acceleration corrected by (1 - rho_f / rho)
 * unless fluid_corrected == 0; the Transform basis of every particle is kept
 * (GrainsBuilderFactory::createCoupledWithFluid + do_before_time_stepping,
 * InterfaceGrains3DBasilisk.cpp:18-45). */
int grains_gpu_create_coupled(const char* xml_path, double fluid_rho, int fluid_corrected, grains_gpu_system** out);
/* One fluid time step: max(min(maxdt, dt_fluid / ndt), mindt) sub-steps
 * (GrainsCoupledWithFluid::Simulation, GrainsCoupledWithFluid.cpp:123-232);
 * *nsub receives their number. */
int grains_gpu_fluid_step(grains_gpu_system* sys, double dt_fluid, int* nsub);
/* GrainsToFluid text (GrainsCoupledWithFluid.cpp:912-1120), tokens joined
 * by single spaces; *buf is allocated with malloc (release with free). */
int grains_gpu_to_fluid(grains_gpu_system* sys, char** buf, int* size);
/* m rows of (vx, vy, vz, wx, wy, wz) by id - 1
 * (GrainsCoupledWithFluid::updateParticlesVelocity, :1126-1181) */
int grains_gpu_set_velocities(grains_gpu_system* sys, const double* vw, size_t m);
/* m rows of the fluid's force and torque on each particle (Fx, Fy, Fz, Tx,
 * Ty, Tz) by id - 1, added to the torsor at every force evaluation from now
 * on, after the contact forces, the force at the centre
 * (GrainsCoupledWithFluid::updateParticlesHydroFT, :1187-1191, ->
 * AppPRSHydroFT::setHydroFT / ComputeForces, AppPRSHydroFT.cpp:27-63; caller
 * FS_Grains3DPlugIn::transferHydroFTtoSolid, FS_Grains3DPlugIn.cc:181-190).
 * m must equal the number of particles. */
int grains_gpu_set_hydro_ft(grains_gpu_system* sys, const double* ft, size_t m);
/* particles and obstacles sent to the fluid (numberOfRBToFluid, :896-905) */
int grains_gpu_count_rigid_bodies(grains_gpu_system* sys, size_t* nparticles, size_t* nobstacles);

/* ---- Discrete adjoint ----------------------------------------------------
 * Reverse-mode derivative of nsteps SecondOrderLeapFrog steps (spheres, Hooke
 * model without rolling resistance, FP64) with respect to the initial
 * positions, velocities, angular velocities and to (kn, en, muc) of one
 * material pair.  No counterpart in the reference, which has no derivative
 * code; the step it differentiates is Grains::Simulation's loop body
 * (Grains.cpp:286-452) with HookeContactForceModel::performForcesCalculus
 * (HookeContactForceModel.cpp:44-114).
 *
 * adjoint_forward advances the system by nsteps steps, keeping the full state
 * every checkpoint_every steps.  adjoint_backward takes the gradient of a
 * scalar J with respect to the final x, v, w (id order, 3 per particle, NULL
 * = zero), rebuilds every segment from its checkpoint, sweeps it in reverse
 * and writes dJ/dx0, dJ/dv0, dJ/dw0 (NULL skipped) and grad_params =
 * {dJ/dkn, dJ/den, dJ/dmuc} of the pair (matA, matB); the system is left at
 * the end state of the forward run.  Memory: 22 doubles per particle per
 * checkpoint plus 9 per particle per step of one segment. */
int grains_gpu_adjoint_forward(grains_gpu_system* sys, int nsteps, int checkpoint_every);
int grains_gpu_adjoint_backward(grains_gpu_system* sys, int matA, int matB, const double* dJ_dx, const double* dJ_dv,
                                const double* dJ_dw, double* grad_x0, double* grad_v0, double* grad_w0,
                                double* grad_params);

/* One Hooke contact evaluated on the host by the derivative code of the
 * adjoint, for checks without a device.  kind 0: body 1 is a sphere, body1 =
 * {x[3], r, v[3], w[3], m}; kind 1: body 1 is a box obstacle, body1 =
 * {L[3], centre[3], basis[9] (row-major rotation)}.  body0 = {x[3], r, v[3],
 * w[3], m} is a sphere; model = {kn, en, etat (-1: from en), muc}.  loads =
 * {F0, T0, F1, T1} (12; zero without contact).  With cot = {F0b, T0b, F1b,
 * T1b} (12) grad receives d(F0b.F0 + T0b.T0 + F1b.F1 + T1b.T1) with respect
 * to body 0's x, v, w and to kn, en, muc (12).  Returns 1 in contact, 0 not,
 * negative on bad arguments. */
int grains_contact_eval(int kind, const double* body0, const double* body1, const double* model, const double* cot,
                        double* loads, double* grad);

#ifdef __cplusplus
}
#endif

#endif /* GRAINS_GPU_H */
