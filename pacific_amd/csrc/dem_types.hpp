// Device-side data layout of one DEM system (structure of arrays in HBM) and
// the parameter blocks the kernels take by value.
#pragma once
#include <cstdint>

namespace dem {

constexpr int kMaxMaterials = 16;

// Hooke-dashpot-Coulomb parameters of one material pair
// (HookeContactForceModel.cpp:11-20, 170-199); beta is computed on the host
// with the host libm so that both sides use the same bits.
struct HookeModel {
  double kn, en, etat, muc, kr, beta;
  int defined;
  int memory;  // HookeMemory: tangential spring kt, rolling friction mur (on when != 0)
  double kt, mur, etarpf, epst;
};

// Contact memory of the HookeMemory model (Component::m_contactMap,
// Component.cpp:634-690): per particle, one entry per partner met by the
// last committed force evaluation -- partner id (obstacles negative),
// cumulative kt * tangential motion, previous normal, cumulative rolling
// spring torque.  Entry j of particle s at [j * N + s]; buffer cur holds the
// last evaluation's entries, the next evaluation writes the other one.
struct ContactMemory {
  int* n[2];
  int* key[2];
  double* kd[2][3];
  double* pn[2][3];
  double* sr[2][3];
  int cap;
  int cur;
};

// Static box obstacle (Box.cpp:27-125, Transform.cpp:71-77, Convex.cpp:29-41)
struct BoxObstacle {
  double B[9];        // basis rows (world <- box)
  double Binv[9];     // transpose (box <- world)
  double centre[3];
  double invorigin[3];
  double extent[3];
  double corners[24];  // Box::setCornersFaces numbering
  double aabb_c[3], aabb_e[3];
  double crust, req;
  int mat, id;
};

// Linked-cell grid.  The GPU grid holds the reference's interior cells; on a
// periodic axis the reference adds one ghost layer on each side, so
// GPU index = reference index - off[a] with off[a] = 1.
struct Grid {
  double lo[3];     // reference local origin (LinkedCell m_LC_local_origin)
  double edge[3];   // cell edge per axis
  double L[3];      // domain size (periodic vector length)
  int n[3];         // GPU cells per axis
  int off[3];       // 1 on periodic axes, 0 otherwise
  int periodic[3];
  int ncells;
  int pad;
};

// Convex shape class of non-spherical particles (ConvexBuilderFactory:
// <Sphere>, <Box>, <Polyhedron>), with the crust-shrink factors of
// RigidBodyWithCrust (RigidBodyWithCrust.cpp:132-168).  Polyhedron vertices
// live in a table shared by all classes (vert_off, nv).
enum ShapeKind { kShapeSphere = 0, kShapeBox = 1, kShapePoly = 2 };
struct ShapeDesc {
  int type, nv, vert_off, pad;
  double radius;      // sphere
  double extent[3];   // box half lengths
  double scaling[3];  // (bbox extent - crust) / bbox extent per axis
  double bext[3];     // bounding-box half extents at the identity transform
  double crust, circR;
  double obb[3];      // OBB half extents (Box/Polyhedron/Sphere::computeBVolume)
};

// <CollisionDetectionAlgorithm> (Grains.cpp:640-770): GJK with Johnson's
// (sv 0) or the signed-volume (sv 1, GJK_SV) sub-distance, relative
// tolerance of the optimality gap, Nesterov acceleration, and the
// bounding-volume pre-test of general pairs (bv 0 none, 1 OBB, 2 OBC).
struct GjkOpts {
  int sv, accel, bv, pad;
  double tol;
};

// Particle state.  x, v, w are double buffered (index cur / 1-cur) because the
// force sweep reads neighbours' positions and velocities while a fused
// kernel writes the next ones.
struct Particles {
  double* x[2][3];
  double* v[2][3];
  double* w[2][3];   // space-fixed angular velocity
  double* wb[3];     // body-fixed angular velocity (integrated state)
  double* q[4];      // quaternion x, y, z, w
  double* F[3];
  double* T[3];
  double* rad;
  double* mass;
  double* inertia;   // sphere: I[0] = I[3] = I[5]
  double* crust;
  double* wfac;      // mass * (1 - rho_f / rho): weight = wfac * g
  double* cfac;      // coupling factor 1 - rho_f / rho
  double* req;       // equivalent-sphere radius (rolling resistance)
  int* id;
  int* mat;
  // Adams-Bashforth n-2 state (allocated only for that scheme)
  double* v_nm2[3];
  double* dUdt_nm2[3];
  double* wb_nm2[3];
  double* dWdt_nm2[3];
  // non-spherical particles (allocated only when the system has any): shape
  // class (-1 = plain sphere), Transform basis rows (double buffered with x),
  // principal moments of inertia, the shape table and its vertex table
  int* shape;
  double* B[2][9];
  double* I3[3];
  const ShapeDesc* shapes;
  const double* verts;
  GjkOpts gjk;  // narrow phase options of the system's general pairs
};

// Cell-sorted view rebuilt every step by the binning kernels.
struct Sorted {
  int* key;        // [N] cell key of particle (storage index)
  int* slot;       // [N] arrival slot within the cell (atomic counter)
  int* count;      // [ncells] particles per cell (zeroed after use)
  int* start;      // [ncells + 1] exclusive prefix of count
  int* block_sums; // scan scratch
  int* tmp_order;  // [N] storage index by (cell, slot)
  int* order;      // [N] storage index by (cell, id)
  int* sid;        // [N] particle id by sorted slot
  int* inv;        // [N] sorted slot by storage index (nullptr: not kept)
  double* posr;    // [4N] x, y, z, radius by sorted slot
};

// Neighbour candidates of each sorted particle for one step, written by
// k_detect and read by the force sweep: entry i of slot k at ent[i * n + k]
// (consecutive slots are consecutive words).  count[k] may exceed cap; such
// a particle is swept in full by the force kernel.
constexpr int kNbrSlotBitsHost = 27;  // slot bits of a list entry
struct NbrList {
  int* count;     // [n]
  unsigned* ent;  // [cap * n]
  int cap, n;
};

// Verlet neighbour list: every pair closer than the contact distance plus
// a skin at the last build, kept until some particle has moved more than
// half the skin since then (each contact of the meantime is then a listed
// pair), so that binning and the neighbour search run only at rebuilds.
// The list is binned on its own grid, whose cells (at least 2 Rmax factor +
// skin wide) hold every such pair within 27 cells; periodic role swaps keep
// following the linked-cell grid.
struct VerletState {
  Grid Gv;
  double* xb[3];  // positions at the last build (storage order)
  double skin;
  double half2;   // (skin / 2)^2
  int stamp;      // > 0: the staleness check sets DevStatus::rebuild = stamp; 0 = no Verlet list; < 0 wrap only
  int pad;
};

// Device-side error / statistics block (one per system)
struct DevStatus {
  int code;          // first error: 1 contact, 2 motion, 3 simulation
  int id0, id1;
  int npairs;        // directed contact records written
  int pair_overflow;
  int n_slip;        // recorded contacts on the Coulomb cap
  int n_recovered;   // ContactError recoveries
  int pad;
  double overlap_max;
  int rebuild;       // Verlet list found stale (the drift's stamp), cleared by the contact filter
  int n_rebuild;     // Verlet list builds
};

enum ErrCode {
  kOk = 0,
  kContactError = 1,
  kMotionError = 2,
  kSimulationError = 3,
  kMemoryOverflow = 4,
  kContactOverflow = 5
};

// Contacts of one step of a system with non-spherical bodies, as the narrow
// phase found them, for the force sweep: record j of sorted slot k at
// [j * n + k] holds the partner (storage index >= 0, or -1 - obstacle
// index), the partner's id, the partner's centre at its periodic image
// (f[0..2]), the contact point (f[3..5]) and the overlap vector (f[6..8]) of
// RigidBodyWithCrust::ClosestPoint, in the order the narrow phase met them.
// A slot whose candidate is not in contact holds ent = kNoFeat (neither a
// storage index nor an obstacle).
constexpr int kNoFeat = -2147483647 - 1;
struct ContactFeat {
  int* count;    // [n] record slots of each sorted slot (may exceed cap: error)
  int* ent;      // [cap * n]
  int* pid;      // [cap * n]
  double* f[9];  // [cap * n] each
  int cap, n;
  // the narrow phase's work list: (sorted slot << 32 | record slot) of
  // every evaluation of the step, and its length
  unsigned long long* items;  // [cap * n]
  int* nitems;
  int item_blocks;  // workgroups of the work-list kernel
};
enum Scheme { kLeapFrog = 0, kFirstOrderExplicit = 1, kSecondOrderExplicit = 2, kAdamsBashforth2 = 3 };

struct StepParams {
  double g[3];
  double dt;
  int n;              // number of particles
  int cur;            // current x/v/w buffer
  int nobs;
  int nmat;
  int record_pairs;
  int pair_cap;
  int scheme;
  int first_step;     // Adams-Bashforth first step (Euler start)
  double* stats;      // contact statistics: 5 partial sums per force block (nullptr: not collected)
  ContactMemory mem;  // HookeMemory tables (mem.cap == 0: none)
  ContactFeat cf;     // contacts found by the narrow phase (systems with non-spherical bodies)
  double* fsum;       // per storage index: sum of the magnitudes of its contact forces (nullptr: not kept)
  const double* hyd;  // fluid force and torque per storage index, component c at hyd[c * n + s] (nullptr: none)
};

}  // namespace dem
