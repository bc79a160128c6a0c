// ===========================================================================
// grains_oracle.cpp -- TEST INFRASTRUCTURE ONLY (the checker, never the product)
//
// A plain FP64 CPU restatement of the serial Grains3D (GRAINS_V1) time step
// for spheres and box obstacles: linked-cell broad phase with the reference's
// cell numbering, push_front cell lists and 13-cell half-shell traversal,
// closed-form sphere-sphere / sphere-box narrow phase, the Hooke (HODC)
// contact force model, torsor accumulation, and the four time integrators.
// Periodic domains are handled the way the reference does it: clone particles
// in ghost cells, role swap when a master enters a ghost cell.
//
// Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
// load this library.  It is compiled with g++ -O2 -ffp-contract=off (see
// oracle/Makefile) so that every expression is evaluated in the order written,
// as in the x86 reference build.
//
// Every function cites the reference file:line it restates (paths relative to
// the PacIFiC repository root).  Pinning of each function (closed forms,
// brute force, invariants) lives in tests/test_oracle_*.py.  Deliberate
// differences from the reference:
//   * obstacles are attached to every cell whose doubled box meets the
//     obstacle AABB (LinkedCell.cpp:1388-1440 refines that test with a
//     box-box GJK); the attached set is a superset, so no contact changes
//     because a cell edge is >= 2 Rmax;
//   * the ContactError recovery path (RigidBodyWithCrust.cpp:534-590) runs
//     GJK on crust-shrunk shapes; for spheres and boxes the GJK fixed point
//     has a closed form, which is used here;
//   * particles are inserted as given, without the insertion overlap test
//     (LinkedCell.cpp:1765-1860 rejects overlapping insertions).
// ===========================================================================
#include <cmath>
#include <cstdint>
#include <cstring>
#include <cstdio>
#include <list>
#include <map>
#include <vector>
#include <algorithm>

namespace {

// GRAINS_V1/Grains/Base/include/Basic.hh:22-29
const double PI = 3.14159265358979323846;
const double EPSILON = 1.0e-10;
const double OBSTACLE_MASS = 1.e20;  // Component.cpp:16

// ----------------------------------------------------------------------------
// Three-vector with the reference's operation order (Group3.cpp:138-320,
// Vector3.cpp:114-160).
struct V3 {
  double c[3];
  V3() { c[0] = c[1] = c[2] = 0.; }
  V3(double x, double y, double z) { c[0] = x; c[1] = y; c[2] = z; }
  double& operator[](int i) { return c[i]; }
  double operator[](int i) const { return c[i]; }
};
inline V3 operator+(V3 const& a, V3 const& b) { return V3(a[0] + b[0], a[1] + b[1], a[2] + b[2]); }
inline V3 operator-(V3 const& a, V3 const& b) { return V3(a[0] - b[0], a[1] - b[1], a[2] - b[2]); }
inline V3 operator-(V3 const& a) { return V3(-a[0], -a[1], -a[2]); }
// Group3 * double (Group3.cpp:160) and double * Group3 (Group3.cpp:316: result *= d)
inline V3 operator*(V3 const& a, double d) { return V3(a[0] * d, a[1] * d, a[2] * d); }
inline V3 operator*(double d, V3 const& a) { return V3(a[0] * d, a[1] * d, a[2] * d); }
inline V3 operator/(V3 const& a, double d) { return V3(a[0] / d, a[1] / d, a[2] / d); }
inline V3& operator+=(V3& a, V3 const& b) { a[0] += b[0]; a[1] += b[1]; a[2] += b[2]; return a; }
inline V3& operator-=(V3& a, V3 const& b) { a[0] -= b[0]; a[1] -= b[1]; a[2] -= b[2]; return a; }
inline V3& operator*=(V3& a, double d) { a[0] *= d; a[1] *= d; a[2] *= d; return a; }
inline V3& operator/=(V3& a, double d) { a[0] /= d; a[1] /= d; a[2] /= d; return a; }
// Group3.cpp:149-153
inline double dot(V3 const& a, V3 const& b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }
// Vector3.cpp:114-118
inline V3 cross(V3 const& a, V3 const& b) {
  return V3(a[1] * b[2] - a[2] * b[1], -a[0] * b[2] + a[2] * b[0], a[0] * b[1] - a[1] * b[0]);
}
// Vector3.cpp:147-152
inline double Norm(V3 const& a) { return std::sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2]); }
// Group3.cpp:390-395
inline void roundv(V3& a, double tol = EPSILON) {
  for (int i = 0; i < 3; ++i) a[i] = std::fabs(a[i]) < tol ? 0. : a[i];
}

// ----------------------------------------------------------------------------
// Quaternion [w, v] (Quaternion.cpp:74-115, 442-470)
struct Q {
  V3 v;
  double w;
  Q() : v(), w(1.) {}
  Q(V3 const& vv, double ww) : v(vv), w(ww) {}
};
inline Q conj(Q const& q) { return Q(-q.v, q.w); }
// Quaternion::operator* (Quaternion.cpp:442-450)
inline Q qmul(Q const& a, Q const& b) {
  double tmp = (a.w * b.w) - dot(a.v, b.v);
  V3 vt = (cross(a.v, b.v) + (a.w * b.v)) + (b.w * a.v);
  return Q(vt, tmp);
}
// operator,(Vector3, Quaternion) = q.multLeftVec(lhs) (Quaternion.cpp:95-100, 468-471)
inline Q vecq(V3 const& lhs, Q const& q) {
  double tmp = dot(-lhs, q.v);
  V3 vt = cross(lhs, q.v) + (lhs * q.w);
  return Q(vt, tmp);
}
// Quaternion::multToVector3 (Quaternion.cpp:108-113)
inline V3 multToVector3(Q const& a, Q const& b) {
  return (cross(a.v, b.v) + (a.w * b.v)) + (b.w * a.v);
}
// Quaternion::setQuaternion(Matrix) (Quaternion.cpp:200-255)
inline Q quatFromMatrix(const double r[3][3]) {
  Q q;
  double den;
  if (r[1][1] > -r[2][2] && r[0][0] > -r[1][1] && r[0][0] > -r[2][2]) {
    den = std::pow(1. + r[0][0] + r[1][1] + r[2][2], 0.5);
    q.w = 0.5 * den;
    q.v[0] = 0.5 * (r[2][1] - r[1][2]) / den;
    q.v[1] = 0.5 * (r[0][2] - r[2][0]) / den;
    q.v[2] = 0.5 * (r[1][0] - r[0][1]) / den;
  } else if (r[1][1] < -r[2][2] && r[0][0] > r[1][1] && r[0][0] > r[2][2]) {
    den = std::pow(1. + r[0][0] - r[1][1] - r[2][2], 0.5);
    q.w = 0.5 * (r[2][1] - r[1][2]) / den;
    q.v[0] = 0.5 * den;
    q.v[1] = 0.5 * (r[0][1] + r[1][0]) / den;
    q.v[2] = 0.5 * (r[2][0] + r[0][2]) / den;
  } else if (r[1][1] > r[2][2] && r[0][0] < r[1][1] && r[0][0] < -r[2][2]) {
    den = std::pow(1. - r[0][0] + r[1][1] - r[2][2], 0.5);
    q.w = 0.5 * (r[0][2] - r[2][0]) / den;
    q.v[0] = 0.5 * (r[0][1] + r[1][0]) / den;
    q.v[1] = 0.5 * den;
    q.v[2] = 0.5 * (r[1][2] + r[2][1]) / den;
  } else if (r[1][1] < r[2][2] && r[0][0] < -r[1][1] && r[0][0] < r[2][2]) {
    den = std::pow(1. - r[0][0] - r[1][1] + r[2][2], 0.5);
    q.w = 0.5 * (r[1][0] - r[0][1]) / den;
    q.v[0] = 0.5 * (r[2][0] + r[0][2]) / den;
    q.v[1] = 0.5 * (r[1][2] + r[2][1]) / den;
    q.v[2] = 0.5 * den;
  }
  return q;
}

// ----------------------------------------------------------------------------
enum Scheme { LEAPFROG = 0, FIRST_ORDER_EXPLICIT = 1, SECOND_ORDER_EXPLICIT = 2, ADAMS_BASHFORTH2 = 3 };
enum ErrCode { OK = 0, CONTACT_ERROR = 1, MOTION_ERROR = 2, SIMULATION_ERROR = 3, CONFIG_ERROR = 4 };

struct HookeParams {  // HookeContactForceModel.cpp:11-20
  double kn, en, etat, muc, kr, beta;
  bool defined;
  // HookeMemoryContactForceModel.cpp:11-29: tangential spring kt, rolling
  // friction mur (on when != 0) with its damping factor etarpf, threshold epst
  bool memory = false;
  double kt = 0., mur = 0., etarpf = 0., epst = 0.;
};

// One entry of Component::m_contactMap (Component.cpp:634-690): touched
// during the current force evaluation, cumulative kt * tangential motion,
// previous normal, cumulative rolling spring torque.
struct ContactMem {
  bool flag;
  V3 kd, pn, sr;
};

struct Part {
  int id;          // 1..N; clones share their master's id
  bool clone;      // tag-2 copy living in a ghost cell
  bool alive;
  int tag, tag_nm1, geoloc, geoloc_nm1, cell, cell_nm1;
  V3 x, v, w, wb;  // centre, translational, space-fixed and body-fixed angular velocity
  Q q;
  V3 F, T, refp;   // torsor (Torsor.cpp:160-190)
  double fsum = 0.;  // sum of the magnitudes of the contact forces applied (error normaliser, not a reference field)
  double tsum = 0.;  // sum of the term scales of hookeForces (error normaliser, not a reference field)
  std::map<int, ContactMem> mem;  // contact memory by partner id (elementary ids are 1)
  double radius, crust, mass, inertia[6], req, density;
  int mat;
  V3 weight;
  // Adams-Bashforth n-2 data (SecondOrderAdamsBashforth.cpp:51-101)
  V3 v_nm2, dUdt_nm2, wb_nm2, dWdt_nm2;
  // convex shape: -1 = sphere of `radius`, else index into System::shapes;
  // basis of the particle's Transform (rotated by composeLeftByRotation
  // every move) and the polyhedron support warm start
  int shape = -1;
  double B[3][3] = {{1., 0., 0.}, {0., 1., 0.}, {0., 0., 1.}};
  int curr = 0;
  double circR = 0.;
};

struct Obs {
  int id;
  V3 extent, centre;
  double B[3][3];      // basis rows
  double Binv[3][3];   // transpose (Transform.cpp:71-77, no scaling)
  V3 invorigin;
  double crust, circR, req;
  int mat;
  V3 corners[8];
  V3 aabb_c, aabb_e;
  V3 F, T;
};

struct Cell {
  int i, j, k, tag, geoloc;
  // side flag per axis: -1 first interior layer, +1 last interior layer, 0 otherwise
  int side[3];
  std::list<int> parts;        // Cell::add is push_front (Cell.cpp:90-93)
  std::vector<int> nbrContact; // push_front order (Cell.cpp:99-102)
  std::vector<int> obstacles;  // push_back with dedup (Cell.cpp:110-124)
  V3 centre;
};

struct PairRec { int self, other; };

#include "convex_oracle.hpp"

struct System {
  // domain (App.cpp:35-58) and grid (LinkedCell.cpp:47-140)
  V3 dom_origin, dom_max, dom_size;
  int periodic[3];
  double cellsize_factor;
  int nbi, nbj, nbk, nb;
  double edge[3];
  V3 lc_origin, lc_max;
  std::vector<Cell> cells;
  std::vector<int> buffer_cells;
  // physics
  V3 gravity;
  double fluid_density;
  double dt, time;
  long step_counter;
  int scheme;
  std::map<int, HookeParams> models;  // key = matA*64+matB, symmetric
  HookeParams default_model;
  // components
  std::vector<Part> pool;
  std::list<int> active;                 // AllComponents::m_ActiveParticles
  std::multimap<int, int> clones;        // id -> pool index
  std::vector<int> master_of_id;         // id -> pool index of the master
  std::vector<int> insertion;            // insertion order of ids
  std::vector<Obs> obstacles;
  std::vector<Shape> shapes;
  CollisionOptions col;                  // <CollisionDetectionAlgorithm>
  bool initialized;
  // diagnostics
  std::vector<PairRec> pairs;
  bool record_pairs;
  double overlap_max, overlap_mean, overlap_count, time_overlap_max, gjk_iter_mean, gjk_count;
  long n_recovered;
  int err;
  int err_id0, err_id1;
  // error messages
  char msg[256];
};

inline double hooke_beta(double en) {  // HookeContactForceModel.cpp:19
  return std::log(en) / std::sqrt(PI * PI + std::log(en) * std::log(en));
}

HookeParams* model_for(System* s, int ma, int mb) {  // ContactBuilderFactory.cpp:22-26
  auto it = s->models.find(ma * 64 + mb);
  if (it != s->models.end() && it->second.defined) return &it->second;
  if (s->default_model.defined) return &s->default_model;
  return nullptr;
}

// Cell::GetCell (Cell.cpp:172-184)
inline void GetCell(System const* s, V3 const& p, int id[3]) {
  id[0] = int(std::floor((p[0] - s->lc_origin[0]) / s->edge[0]));
  id[1] = int(std::floor((p[1] - s->lc_origin[1]) / s->edge[1]));
  id[2] = int(std::floor((p[2] - s->lc_origin[2]) / s->edge[2]));
}
// LinkedCell::getCellNumber (LinkedCell.cpp:1165-1168)
inline int cellNumber(System const* s, int i, int j, int k) { return j * s->nbk * s->nbi + k * s->nbi + i; }
// LinkedCell::getCell (LinkedCell.cpp:1110-1122)
inline int getCell(System const* s, int i, int j, int k) {
  bool valid = -1 < i && i < s->nbi && -1 < j && j < s->nbj && -1 < k && k < s->nbk;
  return valid ? cellNumber(s, i, j, k) : -1;
}

// Periodic vectors attached to a buffer cell (App.cpp:65-240): one per
// non-empty subset of the axes on which the cell lies in a first/last
// interior layer; first layer shifts by +L, last layer by -L.
void periodicVectors(System const* s, Cell const& c, std::vector<V3>& out) {
  out.clear();
  int axes[3], na = 0;
  for (int a = 0; a < 3; ++a)
    if (c.side[a] != 0) axes[na++] = a;
  for (int mask = 1; mask < (1 << na); ++mask) {
    V3 vec;
    for (int b = 0; b < na; ++b)
      if (mask & (1 << b)) {
        int a = axes[b];
        vec[a] = c.side[a] < 0 ? s->dom_size[a] : -s->dom_size[a];
      }
    out.push_back(vec);
  }
}

// LinkedCell::set (LinkedCell.cpp:47-420) for grids with >= 2 interior cells
// per periodic direction.
int buildGrid(System* s, double cellsize) {
  for (int a = 0; a < 3; ++a) {
    double L = s->dom_size[a];
    int n = 1;
    if (cellsize > EPSILON) {
      n = (int)((L + EPSILON) / cellsize);
      if (!n) n = 1;
    }
    s->edge[a] = L / n;
    if (a == 0) s->nbi = n; else if (a == 1) s->nbj = n; else s->nbk = n;
  }
  s->lc_origin = s->dom_origin;
  V3 mv;
  for (int a = 0; a < 3; ++a)
    if (s->periodic[a]) {
      int& n = (a == 0 ? s->nbi : (a == 1 ? s->nbj : s->nbk));
      if (n < 2) {
        snprintf(s->msg, sizeof(s->msg), "periodic direction %d needs >= 2 interior cells", a);
        return CONFIG_ERROR;
      }
      // Point3::Move(-edge) (LinkedCell.cpp:93-115)
      s->lc_origin[a] = s->lc_origin[a] + (-s->edge[a]);
      n += 2;
    }
  s->nb = s->nbi * s->nbj * s->nbk;
  s->lc_max[0] = s->lc_origin[0] + s->nbi * s->edge[0];
  s->lc_max[1] = s->lc_origin[1] + s->nbj * s->edge[1];
  s->lc_max[2] = s->lc_origin[2] + s->nbk * s->edge[2];
  s->cells.assign(s->nb, Cell());
  int nbs[3] = {s->nbi, s->nbj, s->nbk};
  for (int j = 0; j < s->nbj; j++)
    for (int k = 0; k < s->nbk; k++)
      for (int i = 0; i < s->nbi; i++) {
        Cell& c = s->cells[cellNumber(s, i, j, k)];
        c.i = i; c.j = j; c.k = k;
        int ijk[3] = {i, j, k};
        // Cell::setCentre (Cell.cpp:73-81)
        for (int a = 0; a < 3; ++a)
          c.centre[a] = s->lc_origin[a] + ijk[a] * s->edge[a] + s->edge[a] / 2.;
        c.tag = 0;
        bool ghost = false;
        for (int a = 0; a < 3; ++a) {
          c.side[a] = 0;
          if (!s->periodic[a]) continue;
          if (ijk[a] == 0 || ijk[a] == nbs[a] - 1) ghost = true;
          else if (ijk[a] == 1) c.side[a] = -1;
          else if (ijk[a] == nbs[a] - 2) c.side[a] = +1;
        }
        if (ghost) { c.tag = 2; c.side[0] = c.side[1] = c.side[2] = 0; }
        else if (c.side[0] || c.side[1] || c.side[2]) c.tag = 1;
        c.geoloc = c.side[0] + 3 * c.side[1] + 9 * c.side[2];
      }
  // LinkedCell::setCellContactNeighborhood (LinkedCell.cpp:958-985)
  const int cel[3][13] = {{-1, 0, +1, +1, -1, 0, +1, -1, 0, +1, -1, 0, +1},
                          {0, 0, 0, 0, +1, +1, +1, +1, +1, +1, +1, +1, +1},
                          {+1, +1, +1, 0, +1, +1, +1, 0, 0, 0, -1, -1, -1}};
  for (int m = 0; m < s->nb; ++m) {
    Cell& c = s->cells[m];
    std::vector<int> tmp;
    for (int j = 0; j < 13; j++) {
      int nc = getCell(s, c.i + cel[0][j], c.j + cel[1][j], c.k + cel[2][j]);
      if (nc >= 0) tmp.push_back(nc);
    }
    c.nbrContact.assign(tmp.rbegin(), tmp.rend());  // push_front
  }
  s->buffer_cells.clear();
  for (int j = 0; j < s->nbj; j++)
    for (int k = 0; k < s->nbk; k++)
      for (int i = 0; i < s->nbi; i++)
        if (s->cells[cellNumber(s, i, j, k)].tag == 1) s->buffer_cells.push_back(cellNumber(s, i, j, k));
  return OK;
}

// ----------------------------------------------------------------------------
// Narrow phase.  PointContact fields (PointContact.hh): contact point,
// overlap vector, overlap distance.
struct Contact {
  V3 P, ov;
  double d;
  bool any;       // false = PointNoContact
  int nit = -1;   // PointContact iteration count: GJK's, else 1 for a contact point
};

// ClosestPointSPHERESPHERE (RigidBodyWithCrust.cpp:596-640); returns false on
// ContactError.
bool sphereSphere(Part const& A, V3 const& xA, Part const& B, V3 const& xB, Contact& c) {
  V3 AB = xB - xA;
  double distance = Norm(AB) - (A.radius + B.radius);
  if (distance > 0.) { c.any = false; c.d = 1.e20; return true; }
  if (-distance >= A.crust + B.crust) return false;
  c.P = xA + ((A.radius + 0.5 * distance) * AB) / Norm(AB);
  c.ov = distance * (AB / Norm(AB));
  c.d = distance;
  c.any = true;
  return true;
}

// Box::ContactCornerSPHERE / ContactEdgeSPHERE / IntersectionPointSPHERE
// (Box.cpp:466-769).  Returns false when the centre lies inside the box
// (warningSphereCenterInBox defaults to true, Box.hh:143-145).
bool boxSphere(Obs const& o, V3 const& C, double R, double& overlap, V3& cp) {
  const V3& e = o.extent;
  double gx = C[0], gy = C[1], gz = C[2];
  overlap = 1.;
  cp = V3();
  auto corner = [&](int n) {
    V3 dist = C - o.corners[n];
    double nd = Norm(dist);
    if (nd < R) { overlap = nd - R; cp = C - (1. + 0.5 * overlap / nd) * dist; }
  };
  auto edge = [&](int n, int dir) {
    V3 proj = C;
    proj[dir] = o.corners[n][dir];
    V3 dist = proj - o.corners[n];
    double nd = Norm(dist);
    if (nd < R) { overlap = nd - R; cp = C - (1. + 0.5 * overlap / nd) * dist; }
  };
  if (gx > e[0]) {
    if (gy > e[1]) {
      if (gz > e[2]) corner(6);
      else if (gz < -e[2]) corner(5);
      else edge(5, 2);
    } else if (gy < -e[1]) {
      if (gz > e[2]) corner(2);
      else if (gz < -e[2]) corner(1);
      else edge(1, 2);
    } else {
      if (gz > e[2]) edge(6, 1);
      else if (gz < -e[2]) edge(1, 1);
      else {
        double nd = gx - e[0];
        if (nd < R) { overlap = nd - R; cp = V3(e[0] + 0.5 * overlap, C[1], C[2]); }
      }
    }
  } else if (gx < -e[0]) {
    if (gy > e[1]) {
      if (gz > e[2]) corner(7);
      else if (gz < -e[2]) corner(4);
      else edge(4, 2);
    } else if (gy < -e[1]) {
      if (gz > e[2]) corner(3);
      else if (gz < -e[2]) corner(0);
      else edge(0, 2);
    } else {
      if (gz > e[2]) edge(3, 1);
      else if (gz < -e[2]) edge(0, 1);
      else {
        double nd = -e[0] - gx;
        if (nd < R) { overlap = nd - R; cp = V3(-e[0] - 0.5 * overlap, C[1], C[2]); }
      }
    }
  } else {
    if (gy > e[1]) {
      if (gz > e[2]) edge(6, 0);
      else if (gz < -e[2]) edge(4, 0);
      else {
        double nd = gy - e[1];
        if (nd < R) { overlap = nd - R; cp = V3(C[0], e[1] + 0.5 * overlap, C[2]); }
      }
    } else if (gy < -e[1]) {
      if (gz > e[2]) edge(3, 0);
      else if (gz < -e[2]) edge(0, 0);
      else {
        double nd = -e[1] - gy;
        if (nd < R) { overlap = nd - R; cp = V3(C[0], -e[1] - 0.5 * overlap, C[2]); }
      }
    } else {
      if (gz > e[2]) {
        double nd = gz - e[2];
        if (nd < R) { overlap = nd - R; cp = V3(C[0], C[1], e[2] + 0.5 * overlap); }
      } else if (gz < -e[2]) {
        double nd = -e[2] - gz;
        if (nd < R) { overlap = nd - R; cp = V3(C[0], C[1], -e[2] - 0.5 * overlap); }
      } else {
        return false;
      }
    }
  }
  return true;
}

inline V3 applyT(const double B[3][3], V3 const& o, V3 const& p) {  // Transform::operator() (Transform.cpp:310-316)
  return V3(dot(V3(B[0][0], B[0][1], B[0][2]), p) + o[0], dot(V3(B[1][0], B[1][1], B[1][2]), p) + o[1],
            dot(V3(B[2][0], B[2][1], B[2][2]), p) + o[2]);
}

// ClosestPointSPHEREBOX with the sphere as rigid body A (RigidBodyWithCrust.cpp:647-693)
bool sphereBox(Part const& A, V3 const& xA, Obs const& o, Contact& c) {
  double overlap = 0.;
  V3 local = applyT(o.Binv, o.invorigin, xA);
  V3 cpl;
  if (!boxSphere(o, local, A.radius, overlap, cpl)) return false;
  if (overlap < 0.) {
    if (-overlap >= A.crust + o.crust) return false;
    c.P = applyT(o.B, o.centre, cpl);
    V3 AB = c.P - xA;
    c.ov = (overlap / Norm(AB)) * AB;
    c.d = overlap;
    c.any = true;
  } else {
    c.any = false;
    c.d = 1.e20;
  }
  return true;
}

// ContactError recovery (RigidBodyWithCrust.cpp:534-590 + getTransformWithCrust
// :770-798): closed form of GJK between the shapes shrunk by 10 crusts
// (scaling capped below at 0.5).  Returns false if the shrunk shapes still
// intersect (distance < EPSILON), which is fatal in the reference.
double shrinkScale(double extent, double crust) {
  double sc = extent < EPSILON ? 1. : (extent - 10. * crust) / extent;
  return std::max(sc, 0.5);
}
bool recoverSphereSphere(Part const& A, V3 const& xA, Part const& B, V3 const& xB, Contact& c) {
  double ra = A.radius * shrinkScale(A.radius, A.crust);
  double rb = B.radius * shrinkScale(B.radius, B.crust);
  V3 AB = xB - xA;
  double nab = Norm(AB);
  double distance = nab - ra - rb;
  if (distance < EPSILON) return false;
  V3 n = AB / nab;
  V3 pA = xA + ra * n, pB = xB - rb * n;
  V3 ba = pA - pB;
  ba /= Norm(ba);
  c.P = pA / 2.0 + pB / 2.0;
  double imposed = 1. * (A.crust + B.crust);
  c.ov = imposed * ba;
  c.d = -imposed;
  c.any = true;
  return true;
}
bool recoverSphereBox(Part const& A, V3 const& xA, Obs const& o, Contact& c) {
  double ra = A.radius * shrinkScale(A.radius, A.crust);
  V3 es(o.extent[0] * shrinkScale(o.extent[0], o.crust), o.extent[1] * shrinkScale(o.extent[1], o.crust),
        o.extent[2] * shrinkScale(o.extent[2], o.crust));
  V3 local = applyT(o.Binv, o.invorigin, xA);
  V3 q;
  for (int a = 0; a < 3; ++a) q[a] = std::min(std::max(local[a], -es[a]), es[a]);
  V3 dl = local - q;
  double nd = Norm(dl);
  double distance = nd - ra;
  if (distance < EPSILON) return false;
  V3 pAl = local - (ra / nd) * dl;
  V3 pA = applyT(o.B, o.centre, pAl), pB = applyT(o.B, o.centre, q);
  V3 ba = pA - pB;
  ba /= Norm(ba);
  c.P = pA / 2.0 + pB / 2.0;
  double imposed = 1. * (A.crust + o.crust);
  c.ov = imposed * ba;
  c.d = -imposed;
  c.any = true;
  return true;
}

// ----------------------------------------------------------------------------
// General convex pairs: RigidBodyWithCrust::ClosestPoint dispatch and general
// path (RigidBodyWithCrust.cpp:205-354), ClosestPointSPHEREBOX with the box as
// body A (:694-716) and ClosestPoint_ErreurHandling (:534-590).

// A rigid body with crust as ClosestPoint sees it
struct Geo {
  int type;
  ConvexRef c;
  Tf tf;        // Transform without crust
  double crust, circR;
  V3 scaling;   // crust shrink factors
  V3 bext;      // bbox half extents at the identity transform
};

inline int partType(System const* s, Part const& p) { return p.shape < 0 ? SH_SPHERE : s->shapes[p.shape].type; }

Geo geoOf(System* s, Part& p, V3 const& x) {
  Geo g;
  g.type = partType(s, p);
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) g.tf.B[i][j] = p.B[i][j];
  g.tf.o = x;
  g.crust = p.crust;
  g.circR = p.radius;
  if (g.type == SH_SPHERE) {
    double r = p.shape < 0 ? p.radius : s->shapes[p.shape].radius;
    g.c = ConvexRef{SH_SPHERE, r, V3(), nullptr, &p.curr};
    g.bext = V3(r, r, r);
    for (int a = 0; a < 3; ++a) g.scaling[a] = r < EPSILON ? 1. : (r - p.crust) / r;
  } else {
    Shape const& sh = s->shapes[p.shape];
    g.c = ConvexRef{sh.type, sh.radius, sh.extent, &sh, &p.curr};
    g.bext = sh.bext;
    g.scaling = sh.scaling;
  }
  return g;
}

Geo geoOfObs(Obs& o, int* curr) {
  Geo g;
  g.type = SH_BOX;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) g.tf.B[i][j] = o.B[i][j];
  g.tf.o = o.centre;
  g.crust = o.crust;
  g.circR = o.circR;
  g.c = ConvexRef{SH_BOX, 0., o.extent, nullptr, curr};
  g.bext = o.extent;
  for (int a = 0; a < 3; ++a) g.scaling[a] = o.extent[a] < EPSILON ? 1. : (o.extent[a] - o.crust) / o.extent[a];
  return g;
}

// Transform::composeWithScaling (Transform.cpp:108-112, Matrix.cpp:347-358)
Tf scaledTf(Geo const& g, V3 const& sc) {
  Tf t = g.tf;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) t.B[i][j] *= sc[j];
  return t;
}
// getTransformWithCrust(factor, min_scaling) (RigidBodyWithCrust.cpp:770-798)
Tf scaledTfFactor(Geo const& g, double factor, double min_scaling) {
  V3 sc;
  for (int a = 0; a < 3; ++a) {
    sc[a] = g.bext[a] < EPSILON ? 1. : (g.bext[a] - factor * g.crust) / g.bext[a];
    if (min_scaling != 1.) sc[a] = std::max(sc[a], min_scaling);
  }
  return scaledTf(g, sc);
}

// Obstacle-like view of a box particle for Box::IntersectionPointSPHERE
Obs boxView(Geo const& g) {
  Obs o;
  o.id = 0;
  o.extent = g.c.e;
  o.centre = g.tf.o;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      o.B[i][j] = g.tf.B[i][j];
      o.Binv[i][j] = g.tf.B[j][i];
    }
  for (int i = 0; i < 3; ++i) o.invorigin[i] = dot(-V3(o.Binv[i][0], o.Binv[i][1], o.Binv[i][2]), o.centre);
  const V3& e = o.extent;
  o.corners[0] = V3(-e[0], -e[1], -e[2]);
  o.corners[1] = V3(e[0], -e[1], -e[2]);
  o.corners[2] = V3(e[0], -e[1], e[2]);
  o.corners[3] = V3(-e[0], -e[1], e[2]);
  o.corners[4] = V3(-e[0], e[1], -e[2]);
  o.corners[5] = V3(e[0], e[1], -e[2]);
  o.corners[6] = V3(e[0], e[1], e[2]);
  o.corners[7] = V3(-e[0], e[1], e[2]);
  o.crust = g.crust;
  return o;
}

// ClosestPointSPHEREBOX, box = A, sphere = B (RigidBodyWithCrust.cpp:694-716)
bool boxSphereA(Geo const& A, Geo const& B, Contact& c) {
  Obs o = boxView(A);
  double overlap = 0.;
  V3 local = applyT(o.Binv, o.invorigin, B.tf.o);
  V3 cpl;
  if (!boxSphere(o, local, B.circR, overlap, cpl)) return false;
  if (overlap < 0.) {
    if (-overlap >= A.crust + B.crust) return false;
    c.P = applyT(o.B, o.centre, cpl);
    V3 AB = B.tf.o - c.P;
    c.ov = (overlap / Norm(AB)) * AB;
    c.d = overlap;
    c.any = true;
  } else {
    c.any = false;
    c.d = 1.e20;
  }
  return true;
}
// ClosestPointSPHEREBOX, sphere = A, box = B (RigidBodyWithCrust.cpp:661-693)
bool sphereBoxA(Geo const& A, Geo const& B, Contact& c) {
  Obs o = boxView(B);
  double overlap = 0.;
  V3 local = applyT(o.Binv, o.invorigin, A.tf.o);
  V3 cpl;
  if (!boxSphere(o, local, A.circR, overlap, cpl)) return false;
  if (overlap < 0.) {
    if (-overlap >= A.crust + B.crust) return false;
    c.P = applyT(o.B, o.centre, cpl);
    V3 AB = c.P - A.tf.o;
    c.ov = (overlap / Norm(AB)) * AB;
    c.d = overlap;
    c.any = true;
  } else {
    c.any = false;
    c.d = 1.e20;
  }
  return true;
}

long g_gjk_calls = 0, g_gjk_iters = 0, g_gjk_last = 0;

// General path (RigidBodyWithCrust.cpp:277-347); false on ContactError
bool closestPointGJK(Geo const& A, Geo const& B, Contact& c) {
  V3 gcagcb = A.tf.o - B.tf.o;
  if (!(Norm(gcagcb) < A.circR + B.circR)) {
    c.any = false;
    c.d = 1.e20;
    return true;
  }
  // isContactBVolume on the full transforms (RigidBodyWithCrust.cpp:283-284)
  if (g_col.bv && !bvOverlap(g_col.bv, A.c, A.tf, B.c, B.tf)) {
    c.any = false;
    c.d = 1.e20;
    return true;
  }
  Tf a2w = scaledTf(A, A.scaling), b2w = scaledTf(B, B.scaling);
  V3 pa, pb;
  int nb = 0;
  // GJK_SV or Johnson as the input chose (RigidBodyWithCrust.cpp:304-310)
  double distance = g_col.sv ? gjkSVClosestPoints(A.c, B.c, a2w, b2w, pa, pb, nb)
                             : gjkClosestPoints(A.c, B.c, a2w, b2w, pa, pb, nb);
  ++g_gjk_calls;
  g_gjk_iters += nb;
  g_gjk_last = nb;
  c.nit = nb;
  if (distance < EPSILON) return false;
  pa = applyTf(a2w, pa);
  pb = applyTf(b2w, pb);
  V3 ba = pa - pb;
  c.P = pa / 2.0 + pb / 2.0;
  V3 ov = ba / distance;
  roundv(ov);
  ov *= A.crust + B.crust - distance;
  c.ov = ov;
  c.d = distance - (A.crust + B.crust);
  c.any = true;
  return true;
}

// ClosestPoint_ErreurHandling with factor 10 (Particle.cpp:505-507,
// SimpleObstacle.cpp:237-239); false if the shrunk bodies still intersect
bool closestPointRecoverGJK(Geo const& A, Geo const& B, Contact& c) {
  Tf a2w = scaledTfFactor(A, 10., 0.5), b2w = scaledTfFactor(B, 10., 0.5);
  V3 pa, pb;
  int nb = 0;
  double distance = gjkClosestPoints(A.c, B.c, a2w, b2w, pa, pb, nb);
  c.nit = nb;
  if (distance < EPSILON) return false;
  pa = applyTf(a2w, pa);
  pb = applyTf(b2w, pb);
  V3 ba = pa - pb;
  ba /= Norm(ba);
  c.P = pa / 2.0 + pb / 2.0;
  double imposed = 1. * (A.crust + B.crust);
  c.ov = imposed * ba;
  c.d = -imposed;
  c.any = true;
  return true;
}

// ClosestPoint dispatch for pairs with at least one non-sphere body
bool closestPointConvex(Geo const& A, Geo const& B, Contact& c) {
  if (A.type == SH_SPHERE && B.type == SH_BOX) return sphereBoxA(A, B, c);
  if (A.type == SH_BOX && B.type == SH_SPHERE) return boxSphereA(A, B, c);
  return closestPointGJK(A, B, c);
}

// ----------------------------------------------------------------------------
// Hooke force model (HookeContactForceModel.cpp:44-114)
struct Body {  // what performForcesCalculus reads from a Component
  V3 x, v, w;
  double mass, req;
};
inline V3 velAt(Body const& b, V3 const& P) { return b.v + cross(b.w, P - b.x); }  // Particle.cpp:661-665, ParticleKinematics.cpp:229-232

// termscale (optional, not a reference quantity): kn |pen| + (|gamman| +
// |gammat|) |v_rel|, the magnitude of the terms the force is made of, the
// scale that the rounding of a reduced-precision evaluation is relative to.
void hookeForces(HookeParams const& m, Body const& p0, Body const& p1, Contact const& ci, V3& delFN, V3& delFT,
                 V3& delM, double* termscale = nullptr) {
  V3 P = ci.P;
  V3 pen = ci.ov;
  V3 normal(pen);
  normal /= Norm(normal);
  roundv(normal);
  V3 tmpV = velAt(p0, P) - velAt(p1, P);
  V3 v_n = normal * dot(tmpV, normal);
  V3 v_t = tmpV - v_n;
  delFN = m.kn * pen;
  double avmass = 1. / (1. / p0.mass + 1. / p1.mass);
  double gamman = -2. * m.beta * std::sqrt(avmass * m.kn);
  delFN -= gamman * v_n;
  double normFN = Norm(delFN);
  double etat = m.etat == -1. ? -m.beta * std::sqrt(m.kn / avmass) : m.etat;
  double gammat = 2. * etat * avmass;
  delFT = -gammat * v_t;
  if (termscale) *termscale = m.kn * Norm(pen) + (std::fabs(gamman) + std::fabs(gammat)) * Norm(tmpV);
  double fn = m.muc * normFN;
  double ft = Norm(delFT);
  if (fn < ft) {
    double normv_t = Norm(v_t);
    V3 tangent(0., 0., 0.);
    if (normv_t > EPSILON) tangent = -v_t / normv_t;
    delFT = tangent * fn;
  }
  if (m.kr) {
    V3 wrel = p0.w - p1.w;
    double normwrel = Norm(wrel);
    double normwtrel = Norm(cross(p0.w, p0.x - P) - cross(p1.w, p1.x - P));
    if (normwrel > EPSILON) {
      double Req = 1. / (1. / p0.req + 1. / p1.req);
      delM = -(m.kr * Req * normFN * normwtrel / normwrel) * wrel;
    }
  } else
    delM = V3(0., 0., 0.);
}

// Quaternion::setRotFromTwoVectors (Quaternion.cpp:263-286)
Q rotFromTwoVectors(V3 const& u, V3 const& v) {
  double norm_u_norm_v = std::sqrt(dot(u, u) * dot(v, v));
  double real_part = norm_u_norm_v + dot(u, v);
  V3 vect;
  if (real_part < 1.e-6 * norm_u_norm_v) {
    real_part = 0.;
    vect = std::fabs(u[0]) > std::fabs(u[2]) ? V3(-u[1], u[0], 0.) : V3(0., -u[2], u[1]);
  } else {
    vect = cross(u, v);
  }
  Q qq(vect, real_part);
  double nq = std::sqrt(qq.v[0] * qq.v[0] + qq.v[1] * qq.v[1] + qq.v[2] * qq.v[2] + qq.w * qq.w);
  double d = 1. / nq;
  return Q(d * qq.v, d * qq.w);
}
// Quaternion::rotateVector (Quaternion.cpp:293-298)
V3 rotateVector(Q const& q, V3 const& v) {
  double nv = Norm(q.v);
  return ((q.w * q.w - nv * nv) * v + (2. * dot(v, q.v)) * q.v) + (2. * q.w) * cross(q.v, v);
}

// HookeMemoryContactForceModel::computeTangentialVector (:54-65)
V3 memTangent(HookeParams const& m, double n_t, V3 const& ut, V3 const& kdelta) {
  if (Norm(ut) > 0.001) return -ut / Norm(ut);
  V3 a = kdelta + n_t * ut;
  if ((Norm(kdelta) > m.epst || Norm(n_t * ut) > m.epst) && Norm(a) > m.epst) return -a / Norm(a);
  return V3(0., 0., 0.);
}

// HookeMemoryContactForceModel::performForcesCalculus (:71-193): mem0 is
// p0's entry for p1 (created zero when the contact is new); returns the
// entry p1 receives for p0.
ContactMem hookeMemoryForces(HookeParams const& m, Body const& p0, Body const& p1, Contact const& ci, double dt,
                             bool existed, ContactMem& mem0, V3& delFN, V3& delFT, V3& delM) {
  V3 P = ci.P;
  V3 pen = ci.ov;
  V3 normal(pen);
  normal /= Norm(normal);
  roundv(normal);
  V3 tmpV = velAt(p0, P) - velAt(p1, P);
  V3 v_n = normal * dot(tmpV, normal);
  V3 v_t = tmpV - v_n;
  delFN = m.kn * pen;
  double avmass = 1. / (1. / p0.mass + 1. / p1.mass);
  double gamman = -2. * m.beta * std::sqrt(avmass * m.kn);
  delFN -= gamman * v_n;
  double normFN = Norm(delFN);
  double etat = m.etat == -1. ? -m.beta * std::sqrt(m.kn / avmass) : m.etat;
  double gammat = 2. * etat * avmass;
  V3 viscousFT = -gammat * v_t;
  Q qrot(V3(0., 0., 0.), 0.);
  if (existed) {
    qrot = rotFromTwoVectors(mem0.pn, normal);
    mem0.kd = rotateVector(qrot, mem0.kd);
  }
  mem0.kd += (dt * m.kt) * v_t;
  mem0.pn = normal;
  V3 tij = memTangent(m, gammat, v_t, mem0.kd);
  double normFT = Norm(-mem0.kd + viscousFT);
  if (normFT <= m.muc * normFN) {
    delFT = Norm(-mem0.kd + viscousFT) * tij;
  } else {
    if (m.kt) mem0.kd = ((-m.muc) * normFN * tij + viscousFT) / m.kt;
    delFT = (m.muc * normFN) * tij;
  }
  if (m.mur) {
    double Req = 1. / (1. / p0.req + 1. / p1.req);
    double kr = 3. * m.kn * m.mur * m.mur * Req * Req;
    double etar = 3. * gamman * m.mur * m.mur * Req * Req;
    double max_normMk = m.mur * Req * normFN;
    V3 wrel = p0.w - p1.w;
    if (existed) mem0.sr = rotateVector(qrot, mem0.sr);
    mem0.sr -= (kr * dt) * wrel;
    double normMk = Norm(mem0.sr);
    if (normMk > max_normMk) mem0.sr *= max_normMk / normMk;
    delM = mem0.sr - (m.etarpf * etar) * wrel;
  } else {
    delM = V3(0., 0., 0.);
  }
  mem0.flag = true;
  ContactMem m1;
  m1.flag = true;
  m1.kd = -mem0.kd;
  m1.pn = -normal;
  m1.sr = -mem0.sr;
  return m1;
}

// Torsor::addForce(point, f) (Torsor.cpp:184-190)
inline void torsorAdd(V3& F, V3& T, V3 const& refp, V3 const& P, V3 const& f) {
  F += f;
  V3 a = P - refp;
  T += cross(a, f);
}

Body bodyOf(Part const& p) {
  Body b;
  b.x = p.x; b.v = p.v; b.w = p.w; b.mass = p.mass; b.req = p.req;
  return b;
}

// HookeContactForceModel::computeForces (HookeContactForceModel.cpp:121-163)
// for a particle pair.
// Component::getContactMemory with createContact (Component.cpp:634-674)
bool getContactMemory(Part& p, int id, ContactMem*& e) {
  auto it = p.mem.find(id);
  if (it != p.mem.end()) {
    it->second.flag = true;
    e = &it->second;
    return true;
  }
  ContactMem z;
  z.flag = true;
  z.kd = z.pn = z.sr = V3(0., 0., 0.);
  e = &p.mem.emplace(id, z).first->second;
  return false;
}

void forcesPP(System* s, HookeParams const& m, Part& p0, Part& p1, Contact const& ci) {
  V3 fn, ft, dm;
  double coef = 1. / double(1);
  if (m.memory) {
    // HookeMemoryContactForceModel::computeForces (:200-247)
    ContactMem* e0;
    bool existed = getContactMemory(p0, p1.id, e0);
    ContactMem e1 = hookeMemoryForces(m, bodyOf(p0), bodyOf(p1), ci, s->dt, existed, *e0, fn, ft, dm);
    p1.mem[p0.id] = e1;  // Component::addDeplContactInMap (:680-690)
    torsorAdd(p0.F, p0.T, p0.refp, ci.P, coef * (fn + ft));
    p0.fsum += Norm(fn + ft);
    p1.fsum += Norm(fn + ft);
    if (m.mur) p0.T += dm * coef;
    torsorAdd(p1.F, p1.T, p1.refp, ci.P, coef * (-fn - ft));
    if (m.mur) p1.T += -dm * coef;
    return;
  }
  double ts = 0.;
  hookeForces(m, bodyOf(p0), bodyOf(p1), ci, fn, ft, dm, &ts);
  torsorAdd(p0.F, p0.T, p0.refp, ci.P, coef * (fn + ft));
  p0.fsum += Norm(fn + ft);
  p1.fsum += Norm(fn + ft);
  p0.tsum += ts;
  p1.tsum += ts;
  if (m.kr) p0.T += dm * coef;
  torsorAdd(p1.F, p1.T, p1.refp, ci.P, coef * (-fn - ft));
  if (m.kr) p1.T += -dm * coef;
}
// particle (p0) - obstacle (p1); Obstacle::addForce skips clone partners
// (Obstacle.cpp:575-596)
void forcesPO(System* s, HookeParams const& m, Part& p0, Obs& o, Contact const& ci) {
  Body b1;
  b1.x = o.centre; b1.v = V3(); b1.w = V3(); b1.mass = OBSTACLE_MASS; b1.req = o.req;
  V3 fn, ft, dm;
  if (m.memory) {
    // the obstacle's own map is never read back for the particle's force
    ContactMem* e0;
    bool existed = getContactMemory(p0, o.id, e0);
    hookeMemoryForces(m, bodyOf(p0), b1, ci, s->dt, existed, *e0, fn, ft, dm);
  } else {
    double ts = 0.;
    hookeForces(m, bodyOf(p0), b1, ci, fn, ft, dm, &ts);
    p0.tsum += ts;
  }
  const bool rolling = m.memory ? m.mur != 0. : m.kr != 0.;
  double coef = 1. / double(1);
  torsorAdd(p0.F, p0.T, p0.refp, ci.P, coef * (fn + ft));
  p0.fsum += Norm(fn + ft);
  if (rolling) p0.T += dm * coef;
  if (!p0.clone) {
    torsorAdd(o.F, o.T, o.centre, ci.P, coef * (-fn - ft));
    if (rolling) o.T += -dm * coef;
  }
}

void contactStats(System* s, Contact const& c) {  // AppCollision::addToContactsFeatures (AppCollision.cpp:159-189)
  if (c.any && c.d < 0.) {
    double ov = std::fabs(c.d);
    s->overlap_mean = (s->overlap_mean * s->overlap_count + ov) / (s->overlap_count + 1.);
    s->overlap_count += 1.;
    if (ov > s->overlap_max) {
      s->overlap_max = ov;
      s->time_overlap_max = s->time;
    }
  }
  int nit = c.any ? (c.nit >= 0 ? c.nit : 1) : 0;
  if (nit) {
    s->gjk_iter_mean = (s->gjk_iter_mean * s->gjk_count + double(nit)) / (s->gjk_count + 1.);
    s->gjk_count += 1.;
  }
}

// Particle::InterAction (Particle.cpp:477-545) for sphere pairs
int interactPP(System* s, int ia, int ib) {
  Part& A = s->pool[ia];
  Part& B = s->pool[ib];
  Contact c;
  if (A.shape < 0 && B.shape < 0) {
    if (!sphereSphere(A, A.x, B, B.x, c)) {
      if (!recoverSphereSphere(A, A.x, B, B.x, c)) {
        s->err_id0 = A.id; s->err_id1 = B.id;
        return CONTACT_ERROR;
      }
      s->n_recovered++;
    }
  } else {
    Geo ga = geoOf(s, A, A.x), gb = geoOf(s, B, B.x);
    if (!closestPointConvex(ga, gb, c)) {
      if (!closestPointRecoverGJK(ga, gb, c)) {
        s->err_id0 = A.id; s->err_id1 = B.id;
        return CONTACT_ERROR;
      }
      s->n_recovered++;
    }
  }
  contactStats(s, c);
  if (c.any && c.d < 0.) {
    HookeParams* m = model_for(s, A.mat, B.mat);
    if (!m) return CONFIG_ERROR;
    forcesPP(s, *m, A, B, c);
    if (s->record_pairs) {
      if (!A.clone) s->pairs.push_back({A.id, B.id});
      if (!B.clone) s->pairs.push_back({B.id, A.id});
    }
  }
  return OK;
}

// SimpleObstacle::InterAction (SimpleObstacle.cpp:222-271) + BBox::InZone
// (BBox.cpp:166-172)
int interactPO(System* s, int io, int ip) {
  Obs& o = s->obstacles[io];
  Part& A = s->pool[ip];
  bool probable = std::fabs(o.aabb_c[0] - A.x[0]) <= o.aabb_e[0] + A.radius &&
                  std::fabs(o.aabb_c[1] - A.x[1]) <= o.aabb_e[1] + A.radius &&
                  std::fabs(o.aabb_c[2] - A.x[2]) <= o.aabb_e[2] + A.radius;
  Contact c;
  c.any = false;
  c.d = 1.e20;
  if (probable && A.shape < 0) {
    if (!sphereBox(A, A.x, o, c)) {
      if (!recoverSphereBox(A, A.x, o, c)) {
        s->err_id0 = A.id; s->err_id1 = o.id;
        return CONTACT_ERROR;
      }
      s->n_recovered++;
    }
  } else if (probable) {
    int dummy = 0;
    Geo ga = geoOf(s, A, A.x), gb = geoOfObs(o, &dummy);
    if (!closestPointConvex(ga, gb, c)) {
      if (!closestPointRecoverGJK(ga, gb, c)) {
        s->err_id0 = A.id; s->err_id1 = o.id;
        return CONTACT_ERROR;
      }
      s->n_recovered++;
    }
  }
  contactStats(s, c);
  if (c.any && c.d < 0.) {
    HookeParams* m = model_for(s, A.mat, o.mat);
    if (!m) return CONFIG_ERROR;
    forcesPO(s, *m, A, o, c);
    if (s->record_pairs && !A.clone) s->pairs.push_back({A.id, o.id});
  }
  return OK;
}

// LinkedCell::ComputeForces (LinkedCell.cpp:1019-1103)
int computeForcesLC(System* s) {
  std::vector<int> nbrs;
  for (int ip : s->active) {
    Part& ref = s->pool[ip];
    int id[3];
    GetCell(s, ref.x, id);
    int ci = getCell(s, id[0], id[1], id[2]);
    if (ci < 0) { s->err_id0 = ref.id; return SIMULATION_ERROR; }
    Cell& c = s->cells[ci];
    nbrs.clear();
    auto it = std::find(c.parts.begin(), c.parts.end(), ip);
    if (it != c.parts.end()) {
      ++it;
      for (; it != c.parts.end(); ++it) nbrs.push_back(*it);
    }
    for (int nc : c.nbrContact)
      for (int jp : s->cells[nc].parts) nbrs.push_back(jp);
    for (int jp : nbrs) {
      int e = interactPP(s, ip, jp);
      if (e) return e;
    }
    for (int io : c.obstacles) {
      int e = interactPO(s, io, ip);
      if (e) return e;
    }
  }
  return OK;
}

// Grains::computeParticlesForceAndTorque (Grains.cpp:459-482) with
// AllComponents::InitializeForces (AllComponents.cpp:308-320) and
// Particle::InitializeForce (Particle.cpp:1408-1415)
int computeParticlesForceAndTorque(System* s) {
  // AllComponents::setAllContactMapToFalse (AllComponents.cpp:2114-2124)
  for (int ip : s->active)
    for (auto& kv : s->pool[ip].mem) kv.second.flag = false;
  for (int ip : s->active) {
    Part& p = s->pool[ip];
    p.refp = p.x;
    p.F = p.weight;
    p.T = V3(0., 0., 0.);
    p.fsum = 0.;
    p.tsum = 0.;
  }
  for (auto& o : s->obstacles) { o.F = V3(0., 0., 0.); o.T = V3(0., 0., 0.); }
  s->pairs.clear();
  int e = computeForcesLC(s);
  // AllComponents::updateAllContactMaps (:2132-2142): entries not met by
  // this evaluation are dropped
  for (int ip : s->active) {
    auto& mm = s->pool[ip].mem;
    for (auto it = mm.begin(); it != mm.end();) it = it->second.flag ? std::next(it) : mm.erase(it);
  }
  return e;
}

// ----------------------------------------------------------------------------
// Kinematics (ParticleKinematics.cpp:62-128, ParticleKinematicsSphere.cpp:47-55)
inline V3 torqueBF(Part const& p) {
  Q pc = conj(p.q);
  return multToVector3(pc, vecq(p.T, p.q));
}
inline V3 angAccSphere(Part const& p, V3 const& tbf, double cf) {
  V3 d;
  d[0] = tbf[0] / (cf * p.inertia[0]);
  d[1] = tbf[1] / (cf * p.inertia[3]);
  d[2] = tbf[2] / (cf * p.inertia[5]);
  return d;
}
// ParticleKinematics3D::computeAngularAccelerationBodyFixed in predictor mode
// (ParticleKinematics3D.cpp:56-77; Grains::m_predictor_mode defaults to true)
inline V3 angAcc3D(Part const& p, V3 const& tbf, double cf, V3 const& om) {
  const double* I = p.inertia;
  V3 d;
  d[0] = tbf[0] / (cf * I[0]) + om[1] * om[2] * (I[3] - I[5]) / I[0];
  d[1] = tbf[1] / (cf * I[3]) + om[2] * om[0] * (I[5] - I[0]) / I[3];
  d[2] = tbf[2] / (cf * I[5]) + om[0] * om[1] * (I[0] - I[3]) / I[5];
  return d;
}
// KinematicsBuilderFactory::create: spheres get ParticleKinematicsSphere,
// other convexes ParticleKinematics3D
inline V3 angAcc(System const* s, Part const& p, V3 const& tbf, double cf, V3 const& om) {
  if (p.shape >= 0 && s->shapes[p.shape].type != SH_SPHERE) return angAcc3D(p, tbf, cf, om);
  return angAccSphere(p, tbf, cf);
}
inline double couplingFactor(System const* s, Part const& p) {
  double c = 1.;
  if (s->fluid_density != 0.) c -= s->fluid_density / p.density;  // ParticleKinematics.cpp:485-487
  return c;
}
inline V3 spaceFromBody(Q const& q, V3 const& wb) {  // ParticleKinematics.cpp:79-80
  Q pc = conj(q);
  return multToVector3(q, vecq(wb, pc));
}

// SecondOrderLeapFrog SSPRK3 for the body-fixed angular velocity
// (SecondOrderLeapFrog.cpp:57-66)
inline void ssprk3(System const* s, Part const& p, V3 const& tbf, double cf, V3& vrot, double dt) {
  V3 K1 = angAcc(s, p, tbf, cf, vrot);
  K1 *= dt;
  V3 K2 = angAcc(s, p, tbf, cf, vrot + K1);
  K2 *= dt;
  V3 K3 = angAcc(s, p, tbf, cf, vrot + 0.25 * (K1 + K2));
  K3 *= dt;
  vrot += (K1 + K2 + 4. * K3) / 6.;
}

// ParticleKinematics::Move + time integrator Move (returns |dx|)
double kinMove(System* s, Part& p, double dtv, double dtx) {
  double cf = couplingFactor(s, p);
  V3 tbf = torqueBF(p);
  V3 transMotion, meanW;
  V3& v = p.v;
  V3& vrot = p.wb;
  switch (s->scheme) {
    case LEAPFROG:  // SecondOrderLeapFrog.cpp:44-71
      v += (p.F / (p.mass * cf)) * dtv;
      transMotion = v * dtx;
      ssprk3(s, p, tbf, cf, vrot, dtv);
      meanW = vrot;
      break;
    case FIRST_ORDER_EXPLICIT: {  // FirstOrderExplicit.cpp:44-61
      transMotion = v * dtx;
      v += (p.F / (p.mass * cf)) * dtv;
      V3 d = angAcc(s, p, tbf, cf, vrot);
      meanW = vrot;
      vrot += d * dtv;
      break;
    }
    case SECOND_ORDER_EXPLICIT: {  // SecondOrderExplicit.cpp:44-63
      V3 dUdt = p.F / (p.mass * cf);
      transMotion = v * dtx + 0.5 * dUdt * dtx * dtx;
      v += dUdt * dtv;
      V3 d = angAcc(s, p, tbf, cf, vrot);
      meanW = vrot + 0.5 * d * dtx;
      vrot += d * dtv;
      break;
    }
    case ADAMS_BASHFORTH2: {  // SecondOrderAdamsBashforth.cpp:51-101
      if (s->step_counter == 1) {
        V3 dUdt = p.F / (p.mass * cf);
        transMotion = v * dtx;
        p.v_nm2 = v;
        v += dUdt * dtv;
        p.dUdt_nm2 = dUdt;
        V3 d = angAcc(s, p, tbf, cf, vrot);
        meanW = vrot;
        p.wb_nm2 = vrot;
        vrot += d * dtv;
        p.dWdt_nm2 = d;
      } else {
        V3 dUdt = p.F / (p.mass * cf);
        transMotion = 0.5 * (3. * v - p.v_nm2) * dtx;
        p.v_nm2 = v;
        v += 0.5 * (3. * dUdt - p.dUdt_nm2) * dtv;
        p.dUdt_nm2 = dUdt;
        V3 d = angAcc(s, p, tbf, cf, vrot);
        meanW = 0.5 * (3. * vrot - p.wb_nm2);
        p.wb_nm2 = vrot;
        vrot += 0.5 * (3. * d - p.dWdt_nm2) * dtv;
        p.dWdt_nm2 = d;
      }
      break;
    }
  }
  p.w = spaceFromBody(p.q, vrot);
  meanW = spaceFromBody(p.q, meanW);
  p.x = p.x + transMotion;  // Transform::composeLeftByTranslation (Transform.cpp:300-304)
  double nOmega = Norm(meanW);
  Q qdt;
  if (nOmega > EPSILON) {
    double c = std::cos(nOmega * dtx / 2.);
    double sn = std::sin(nOmega * dtx / 2.);
    V3 t = (sn * 1. / nOmega) * meanW;
    qdt = Q(t, c);
  } else
    qdt = Q(V3(0., 0., 0.), 1.);
  p.q = qmul(qdt, p.q);
  composeLeftByRotation(p.B, qdt);  // Particle::Rotate -> RigidBody::Rotate (RigidBody.cpp:268-273)
  return Norm(transMotion);
}

// ParticleKinematics::advanceVelocity (ParticleKinematics.cpp:111-128) with
// SecondOrderLeapFrog::advanceVelocity (SecondOrderLeapFrog.cpp:74-95)
void kinAdvance(System* s, Part& p, double dtv) {
  double cf = couplingFactor(s, p);
  V3 tbf = torqueBF(p);
  p.v += (p.F / (p.mass * cf)) * dtv;
  ssprk3(s, p, tbf, cf, p.wb, dtv);
  p.w = spaceFromBody(p.q, p.wb);
}

// ----------------------------------------------------------------------------
// Linking (LinkedCell.cpp:1352-1366, 1452-1550)
int linkParticle(System* s, int ip) {
  Part& p = s->pool[ip];
  int id[3];
  GetCell(s, p.x, id);
  int c = getCell(s, id[0], id[1], id[2]);
  if (c < 0) { s->err_id0 = p.id; return SIMULATION_ERROR; }
  p.cell = p.cell_nm1 = c;
  p.tag = p.tag_nm1 = s->cells[c].tag;
  p.geoloc = p.geoloc_nm1 = s->cells[c].geoloc;
  auto& L = s->cells[c].parts;
  if (std::find(L.begin(), L.end(), ip) == L.end()) L.push_front(ip);
  return OK;
}

int linkUpdate(System* s) {
  for (int ip : s->active) {
    Part& p = s->pool[ip];
    p.cell_nm1 = p.cell;
    p.tag_nm1 = p.tag;
    p.geoloc_nm1 = p.geoloc;
    int id[3];
    GetCell(s, p.x, id);
    int c = getCell(s, id[0], id[1], id[2]);
    if (c < 0) { s->err_id0 = p.id; return SIMULATION_ERROR; }
    if (c != p.cell_nm1) {
      s->cells[c].parts.push_front(ip);
      s->cells[p.cell_nm1].parts.remove(ip);
    }
    p.cell = c;
    p.tag = s->cells[c].tag;
    p.geoloc = s->cells[c].geoloc;
  }
  return OK;
}

// Particle::createCloneCopy + Translate (LinkedCell.cpp:1830-1850)
int makeClone(System* s, int im, V3 const& vec) {
  Part c = s->pool[im];
  c.clone = true;
  c.x = c.x + vec;
  // setAngularVelocity recomputes the body-fixed velocity (ParticleKinematics.cpp:197-203)
  c.wb = multToVector3(conj(c.q), vecq(c.w, c.q));
  s->pool.push_back(c);
  int ic = (int)s->pool.size() - 1;
  return ic;
}

inline bool isInLC(System const* s, V3 const& p) {  // LinkedCell.cpp:1650-1660
  return !(p[0] < s->lc_origin[0] || p[0] > s->lc_max[0] || p[1] < s->lc_origin[1] || p[1] > s->lc_max[1] ||
           p[2] < s->lc_origin[2] || p[2] > s->lc_max[2]);
}

// Copies the moved master onto its clone (LinkedCell.cpp:2004-2090)
void copyToClone(System* s, Part const& m, Part& c, V3 const& vec) {
  (void)s;
  c.x = m.x + vec;
  c.v = m.v;
  c.q = m.q;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) c.B[i][j] = m.B[i][j];
  c.w = m.w;
  c.wb = multToVector3(conj(c.q), vecq(c.w, c.q));
}

void updateDestroyPeriodicClones(System* s) {  // LinkedCell.cpp:2004-2128
  std::vector<V3> vecs;
  for (int cb : s->buffer_cells) {
    Cell& cell = s->cells[cb];
    periodicVectors(s, cell, vecs);
    for (int ip : cell.parts) {
      Part& m = s->pool[ip];
      auto range = s->clones.equal_range(m.id);
      size_t ncid = std::distance(range.first, range.second);
      if (ncid == 1) {
        copyToClone(s, m, s->pool[range.first->second], vecs[0]);
      } else if (ncid > 1) {
        for (size_t i = 0; i < vecs.size(); ++i) {
          V3 target = m.x + vecs[i];
          for (auto it = range.first; it != range.second; ++it) {
            Part& c = s->pool[it->second];
            if (Norm(c.x - target) < 2. * m.crust) { copyToClone(s, m, c, vecs[i]); break; }
          }
        }
      }
    }
  }
  for (auto it = s->clones.begin(); it != s->clones.end();) {
    Part& c = s->pool[it->second];
    if (isInLC(s, c.x)) { ++it; continue; }
    s->cells[c.cell].parts.remove(it->second);
    s->active.remove(it->second);
    c.alive = false;
    it = s->clones.erase(it);
  }
}

void createDestroyPeriodicClones(System* s) {  // LinkedCell.cpp:2133-2310
  std::vector<V3> vecs;
  std::vector<int> snapshot(s->active.begin(), s->active.end());
  for (int ip : snapshot) {
    Part& p = s->pool[ip];
    switch (p.tag_nm1) {
      case 0:
        if (p.tag == 1) {
          periodicVectors(s, s->cells[p.cell], vecs);
          for (auto& vec : vecs) {
            int ic = makeClone(s, ip, vec);
            linkParticle(s, ic);
            s->clones.insert({s->pool[ip].id, ic});
            s->active.push_back(ic);
          }
        }
        break;
      case 1:
        if (p.tag == 2) {
          // the master entered a ghost cell: it becomes a clone
          s->pool[ip].clone = true;
          s->clones.insert({p.id, ip});
        } else if (p.tag == 1 && p.geoloc != p.geoloc_nm1) {
          periodicVectors(s, s->cells[p.cell], vecs);
          auto range = s->clones.equal_range(p.id);
          for (auto& vec : vecs) {
            V3 target = s->pool[ip].x + vec;
            bool found = false;
            for (auto it = range.first; it != range.second && !found; ++it)
              if (Norm(s->pool[it->second].x - target) < 2. * s->pool[ip].crust) found = true;
            if (!found) {
              int ic = makeClone(s, ip, vec);
              linkParticle(s, ic);
              s->clones.insert({s->pool[ip].id, ic});
              s->active.push_back(ic);
              range = s->clones.equal_range(s->pool[ip].id);
            }
          }
        }
        break;
      default:
        if (p.tag == 1) {
          // a clone entered the buffer zone: it becomes the master
          auto range = s->clones.equal_range(p.id);
          for (auto it = range.first; it != range.second; ++it)
            if (it->second == ip) { s->clones.erase(it); break; }
          s->pool[ip].clone = false;
          s->master_of_id[p.id] = ip;
        }
        break;
    }
  }
}

// Grains::moveParticlesAndObstacles (Grains.cpp:489-527)
int moveParticles(System* s, double dtv, double dtx) {
  for (int ip : s->active) {
    Part& p = s->pool[ip];
    if (p.clone) continue;
    double depl = kinMove(s, p, dtv, dtx);
    if (depl > p.crust) { s->err_id0 = p.id; return MOTION_ERROR; }  // Particle.cpp:444-456
  }
  bool per = s->periodic[0] || s->periodic[1] || s->periodic[2];
  if (per) updateDestroyPeriodicClones(s);
  int e = linkUpdate(s);
  if (e) return e;
  if (per) createDestroyPeriodicClones(s);
  return OK;
}

// One pass of the Grains::Simulation loop body (Grains.cpp:286-452)
int oneStep(System* s) {
  s->time += s->dt;
  s->step_counter++;
  int e;
  if (s->scheme == LEAPFROG) {
    e = moveParticles(s, 0.5 * s->dt, s->dt);
    if (e) return e;
    e = computeParticlesForceAndTorque(s);
    if (e) return e;
    for (int ip : s->active) {
      Part& p = s->pool[ip];
      if (!p.clone) kinAdvance(s, p, 0.5 * s->dt);
    }
  } else {
    e = computeParticlesForceAndTorque(s);
    if (e) return e;
    e = moveParticles(s, s->dt, s->dt);
    if (e) return e;
  }
  return OK;
}

}  // namespace

// ===========================================================================
// C API for the tests (ctypes)
// ===========================================================================
extern "C" {

void* orc_new(const double* origin, const double* dmax, const int* periodic, double cellsize_factor,
              const double* gravity, double fluid_density, double dt, int scheme) {
  System* s = new System();
  for (int a = 0; a < 3; ++a) {
    s->dom_origin[a] = origin[a];
    s->dom_max[a] = dmax[a];
    s->dom_size[a] = dmax[a] - origin[a];
    s->periodic[a] = periodic[a] ? 1 : 0;
    s->gravity[a] = gravity[a];
  }
  s->cellsize_factor = cellsize_factor < 1. ? 1. : cellsize_factor;
  s->fluid_density = fluid_density;
  s->dt = dt;
  s->time = 0.;
  s->step_counter = 0;
  s->scheme = scheme;
  s->default_model.defined = false;
  s->initialized = false;
  s->record_pairs = true;
  s->overlap_max = s->overlap_mean = s->overlap_count = 0.;
  s->time_overlap_max = s->gjk_iter_mean = s->gjk_count = 0.;
  s->n_recovered = 0;
  s->err = 0;
  s->err_id0 = s->err_id1 = 0;
  s->msg[0] = 0;
  s->master_of_id.assign(1, -1);
  return s;
}

void orc_free(void* h) { delete (System*)h; }

// material indices are small integers; matB < 0 selects the Default model
// HookeMemoryContactForceModel::defineParameters (:253-307): p = {kn, en,
// etat, muc, kt, mur, etarpf, epst}
int orc_set_memory_contact_model(void* h, int matA, int matB, const double* p) {
  System* s = (System*)h;
  HookeParams m;
  m.kn = p[0]; m.en = p[1]; m.etat = p[2]; m.muc = p[3]; m.kr = 0.;
  m.memory = true;
  m.kt = p[4]; m.mur = p[5]; m.etarpf = p[6]; m.epst = p[7];
  m.beta = hooke_beta(m.en);
  m.defined = true;
  if (matA < 0 || matB < 0) { s->default_model = m; return OK; }
  s->models[matA * 64 + matB] = m;
  s->models[matB * 64 + matA] = m;
  return OK;
}

int orc_set_contact_model(void* h, int matA, int matB, double kn, double en, double etat, double muc, double kr) {
  System* s = (System*)h;
  HookeParams m;
  m.kn = kn; m.en = en; m.etat = etat; m.muc = muc; m.kr = kr;
  m.beta = hooke_beta(en);
  m.defined = true;
  if (matA < 0 || matB < 0) { s->default_model = m; return OK; }
  s->models[matA * 64 + matB] = m;
  s->models[matB * 64 + matA] = m;
  return OK;
}

// Particle construction for a sphere (Particle.cpp:150-166, Sphere.cpp:58-64,
// Sphere.cpp:152-155, Component.cpp:205-208, Particle.cpp:413-418).  Spheres are
// added in insertion order and get ids 1, 2, ...; q is [w, x, y, z] = identity
// unless a basis is given.
int orc_add_sphere(void* h, double radius, double density, double crust, int material, const double* x,
                   const double* v, const double* w, const double* basis9) {
  System* s = (System*)h;
  Part p;
  p.tag = p.tag_nm1 = p.geoloc = p.geoloc_nm1 = p.cell = p.cell_nm1 = 0;
  p.id = (int)s->insertion.size() + 1;
  p.clone = false;
  p.alive = true;
  p.radius = radius;
  p.crust = crust;
  p.density = density;
  p.mat = material;
  double volume = 4.0 * PI * std::pow(radius, 3.) / 3.0;
  p.mass = density * volume;
  double I = 8.0 * PI * std::pow(radius, 5.) / 15.0;
  p.inertia[0] = p.inertia[3] = p.inertia[5] = I;
  p.inertia[1] = p.inertia[2] = p.inertia[4] = 0.;
  for (int i = 0; i < 6; ++i) p.inertia[i] *= density;
  p.req = std::pow((0.75 / PI) * volume, 1. / 3.);
  p.weight = p.mass * (1. - s->fluid_density / density) * s->gravity;
  p.x = V3(x[0], x[1], x[2]);
  if (basis9) {
    double r[3][3];
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) r[i][j] = p.B[i][j] = basis9[3 * i + j];
    p.q = quatFromMatrix(r);
  } else
    p.q = Q(V3(0., 0., 0.), 1.);
  p.circR = radius;
  p.v = V3(v[0], v[1], v[2]);
  p.w = V3(w[0], w[1], w[2]);
  p.wb = multToVector3(conj(p.q), vecq(p.w, p.q));  // ParticleKinematics::setAngularVelocity
  s->pool.push_back(p);
  s->insertion.push_back(p.id);
  s->master_of_id.push_back((int)s->pool.size() - 1);
  return p.id;
}

// Convex shape class (ConvexBuilderFactory: <Sphere Radius>, <Box LX LY LZ>,
// <Polyhedron Name> with a vertex / face file) with its crust thickness;
// returns the shape index.  type 0 sphere (radius), 1 box (edge lengths L),
// 2 polyhedron (nv vertices, nf faces given as face_sizes[nf] and the
// concatenated vertex indices).
int orc_add_shape(void* h, int type, double radius, const double* L, int nv, const double* verts, int nf,
                  const int* face_sizes, const int* face_idx, double crust) {
  System* s = (System*)h;
  Shape sh;
  sh.type = type;
  sh.radius = radius;
  sh.crust = crust;
  sh.extent = V3();
  for (int i = 0; i < 6; ++i) sh.inertia[i] = 0.;
  if (type == SH_SPHERE) {
    sh.volume = 4.0 * PI * std::pow(radius, 3.) / 3.0;
    sh.inertia[0] = sh.inertia[3] = sh.inertia[5] = 8.0 * PI * std::pow(radius, 5.) / 15.0;
    sh.circR = radius;
    sh.start_vertex = 0;
  } else if (type == SH_BOX) {
    sh.extent = V3(L[0] / 2., L[1] / 2., L[2] / 2.);
    buildBoxShape(sh);
  } else if (type == SH_POLY) {
    for (int i = 0; i < nv; ++i) sh.verts.push_back(V3(verts[3 * i], verts[3 * i + 1], verts[3 * i + 2]));
    int off = 0;
    for (int f = 0; f < nf; ++f) {
      std::vector<int> face(face_idx + off, face_idx + off + face_sizes[f]);
      off += face_sizes[f];
      sh.faces.push_back(face);
    }
    buildPolyhedron(sh);
  } else {
    return -1;
  }
  finishShape(sh);
  s->shapes.push_back(sh);
  return (int)s->shapes.size() - 1;
}

// A particle of a shape class (Particle(DOMNode*), Particle.cpp:130-166):
// mass = density * volume, inertia = density * shape inertia; basis9 is the
// row-major rotation of <AngularPosition> (NULL = identity), from which the
// quaternion follows (Quaternion::setQuaternion(Matrix)).
int orc_add_particle(void* h, int shape, double density, int material, const double* x, const double* v,
                     const double* w, const double* basis9) {
  System* s = (System*)h;
  if (shape < 0 || shape >= (int)s->shapes.size()) return -1;
  Shape const& sh = s->shapes[shape];
  if (sh.type == SH_SPHERE) {
    int id = orc_add_sphere(h, sh.radius, density, sh.crust, material, x, v, w, basis9);
    return id;
  }
  Part p;
  p.tag = p.tag_nm1 = p.geoloc = p.geoloc_nm1 = p.cell = p.cell_nm1 = 0;
  p.id = (int)s->insertion.size() + 1;
  p.clone = false;
  p.alive = true;
  p.shape = shape;
  p.radius = sh.circR;
  p.circR = sh.circR;
  p.crust = sh.crust;
  p.density = density;
  p.mat = material;
  p.mass = density * sh.volume;
  for (int i = 0; i < 6; ++i) p.inertia[i] = sh.inertia[i] * density;
  p.req = std::pow((0.75 / PI) * sh.volume, 1. / 3.);
  p.weight = p.mass * (1. - s->fluid_density / density) * s->gravity;
  p.x = V3(x[0], x[1], x[2]);
  if (basis9) {
    double r[3][3];
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) r[i][j] = p.B[i][j] = basis9[3 * i + j];
    p.q = quatFromMatrix(r);
  } else
    p.q = Q(V3(0., 0., 0.), 1.);
  p.curr = sh.start_vertex;
  p.v = V3(v[0], v[1], v[2]);
  p.w = V3(w[0], w[1], w[2]);
  p.wb = multToVector3(conj(p.q), vecq(p.w, p.q));
  s->pool.push_back(p);
  s->insertion.push_back(p.id);
  s->master_of_id.push_back((int)s->pool.size() - 1);
  return p.id;
}

// Shape properties: out = {volume, circumscribed radius, inertia[6],
// scaling[3], bbox extent[3]} per unit density
int orc_shape_info(void* h, int shape, double* out) {
  System* s = (System*)h;
  if (shape < 0 || shape >= (int)s->shapes.size()) return -1;
  Shape const& sh = s->shapes[shape];
  out[0] = sh.volume;
  out[1] = sh.circR;
  for (int i = 0; i < 6; ++i) out[2 + i] = sh.inertia[i];
  for (int a = 0; a < 3; ++a) out[8 + a] = sh.scaling[a];
  for (int a = 0; a < 3; ++a) out[11 + a] = sh.bext[a];
  return 0;
}

// GJK distance between two shapes placed by (basis9, origin3) each, scaled
// by their crust shrink factors when `crust` != 0: returns the distance and
// writes the world witness points pa, pb and the iteration count
double orc_gjk(void* h, int shapeA, const double* BA, const double* oA, int shapeB, const double* BB,
               const double* oB, int crust, double* pa, double* pb, int* iters) {
  System* s = (System*)h;
  g_col = s->col;
  Shape const& sa = s->shapes[shapeA];
  Shape const& sb = s->shapes[shapeB];
  int ca = sa.start_vertex, cb = sb.start_vertex;
  ConvexRef A{sa.type, sa.radius, sa.extent, &sa, &ca}, B{sb.type, sb.radius, sb.extent, &sb, &cb};
  Tf ta, tb;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      ta.B[i][j] = BA[3 * i + j] * (crust ? sa.scaling[j] : 1.);
      tb.B[i][j] = BB[3 * i + j] * (crust ? sb.scaling[j] : 1.);
    }
  ta.o = V3(oA[0], oA[1], oA[2]);
  tb.o = V3(oB[0], oB[1], oB[2]);
  V3 p, q;
  int it = 0;
  double d = gjkClosestPoints(A, B, ta, tb, p, q, it);
  p = applyTf(ta, p);
  q = applyTf(tb, q);
  for (int a = 0; a < 3; ++a) {
    pa[a] = p[a];
    pb[a] = q[a];
  }
  *iters = it;
  return d;
}

// The oracle's narrow phase for two bodies of shape classes placed at
// centres x with bases B: out = {P[3], ov[3], d, status} with status 1
// features set, 0 no contact, -1 ContactError, 2 recovered.  Spheres of a
// sphere class go through the sphere-sphere closed form as particles added
// with orc_add_sphere do.
int orc_closest_point(void* h, int shapeA, const double* xA, const double* BA, int shapeB, const double* xB,
                      const double* BB, double* out) {
  System* s = (System*)h;
  g_col = s->col;
  auto mk = [&](int sh, const double* x, const double* B) {
    Shape const& d = s->shapes[sh];
    Part p;
    p.shape = d.type == SH_SPHERE ? -1 : sh;
    p.radius = d.type == SH_SPHERE ? d.radius : d.circR;
    p.circR = d.circR;
    p.crust = d.crust;
    p.curr = d.start_vertex;
    p.x = V3(x[0], x[1], x[2]);
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) p.B[i][j] = B[3 * i + j];
    return p;
  };
  Part A = mk(shapeA, xA, BA), B = mk(shapeB, xB, BB);
  Contact c;
  c.any = false;
  c.d = 1.e20;
  int st;
  if (A.shape < 0 && B.shape < 0) {
    st = sphereSphere(A, A.x, B, B.x, c) ? 1 : -1;
    if (st < 0 && recoverSphereSphere(A, A.x, B, B.x, c)) st = 2;
  } else {
    Geo ga = geoOf(s, A, A.x), gb = geoOf(s, B, B.x);
    st = closestPointConvex(ga, gb, c) ? 1 : -1;
    if (st < 0 && closestPointRecoverGJK(ga, gb, c)) st = 2;
  }
  if (st == 1 && !c.any) st = 0;
  for (int a = 0; a < 3; ++a) {
    out[a] = c.any ? c.P[a] : 0.;
    out[3 + a] = c.any ? c.ov[a] : 0.;
  }
  out[6] = c.any ? c.d : 1.e20;
  out[7] = st;
  return 0;
}

// Particle bases (id order, 9 per particle, row-major)
void orc_get_basis(void* h, double* B) {
  System* s = (System*)h;
  for (size_t i = 0; i < s->insertion.size(); ++i) {
    Part const& p = s->pool[s->master_of_id[s->insertion[i]]];
    for (int r = 0; r < 3; ++r)
      for (int c = 0; c < 3; ++c) B[9 * i + 3 * r + c] = p.B[r][c];
  }
}

// Overwrites the particle bases (id order, 9 per particle); clones follow
// their master on the next step
void orc_set_basis(void* h, const double* B) {
  System* s = (System*)h;
  for (size_t i = 0; i < s->insertion.size(); ++i) {
    Part& p = s->pool[s->master_of_id[s->insertion[i]]];
    for (int r = 0; r < 3; ++r)
      for (int c = 0; c < 3; ++c) p.B[r][c] = B[9 * i + 3 * r + c];
  }
}

// GJK statistics: calls and total iterations since the library was loaded
void orc_gjk_stats(long* calls, long* iters) {
  *calls = g_gjk_calls;
  *iters = g_gjk_iters;
}
// iterations of the last GJK call
long orc_gjk_last() { return g_gjk_last; }

// n spheres at once (same construction as orc_add_sphere, no basis)
int orc_add_spheres(void* h, int n, const double* radius, const double* density, const double* crust,
                    const int* material, const double* x, const double* v, const double* w) {
  for (int i = 0; i < n; ++i)
    orc_add_sphere(h, radius[i], density[i], crust[i], material[i], x + 3 * i, v + 3 * i, w + 3 * i, nullptr);
  return n;
}

// Box obstacle (Box.cpp:27-33, 92-125; SimpleObstacle bbox Convex.cpp:29-41,
// BBox.cpp:235-239); obstacles get ids -1, -2, ...
int orc_add_box(void* h, const double* L, const double* centre, const double* basis9, double crust, int material) {
  System* s = (System*)h;
  Obs o;
  o.id = -(int)s->obstacles.size() - 1;
  o.extent = V3(std::fabs(L[0]) / 2., std::fabs(L[1]) / 2., std::fabs(L[2]) / 2.);
  o.centre = V3(centre[0], centre[1], centre[2]);
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      o.B[i][j] = basis9 ? basis9[3 * i + j] : (i == j ? 1. : 0.);
    }
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) o.Binv[i][j] = o.B[j][i];
  for (int i = 0; i < 3; ++i) o.invorigin[i] = dot(-V3(o.Binv[i][0], o.Binv[i][1], o.Binv[i][2]), o.centre);
  const V3& e = o.extent;
  o.corners[0] = V3(-e[0], -e[1], -e[2]);
  o.corners[1] = V3(e[0], -e[1], -e[2]);
  o.corners[2] = V3(e[0], -e[1], e[2]);
  o.corners[3] = V3(-e[0], -e[1], e[2]);
  o.corners[4] = V3(-e[0], e[1], -e[2]);
  o.corners[5] = V3(e[0], e[1], -e[2]);
  o.corners[6] = V3(e[0], e[1], e[2]);
  o.corners[7] = V3(-e[0], e[1], e[2]);
  o.crust = crust;
  o.mat = material;
  o.circR = Norm(o.extent);
  o.req = std::pow((0.75 / PI) * (8.0 * e[0] * e[1] * e[2]), 1. / 3.);
  // Convex::bbox with Box::support (Box.cpp:307-316)
  auto support = [&](V3 const& d) {
    double n = Norm(d);
    if (n < EPSILON) return V3();
    return V3(d[0] < 0. ? -e[0] : e[0], d[1] < 0. ? -e[1] : e[1], d[2] < 0. ? -e[2] : e[2]);
  };
  V3 mn, mx;
  for (int a = 0; a < 3; ++a) {
    V3 row(o.B[a][0], o.B[a][1], o.B[a][2]);
    mn[a] = o.centre[a] + dot(row, support(-row));
    mx[a] = o.centre[a] + dot(row, support(row));
  }
  o.aabb_e = (mx - mn) / 2.;
  o.aabb_c = mn + o.aabb_e;
  o.F = V3(); o.T = V3();
  s->obstacles.push_back(o);
  return o.id;
}

// Builds the grid, links particles and obstacles, creates periodic clones
// and primes the LeapFrog forces (Grains.cpp:95-179, 1010-1033, 1965-2010).
// The bounding-volume pre-test (bv 1 OBB, 2 OBC) of two shape classes
// placed at origins o with bases B (rows); 1 when the volumes meet
int orc_bv_overlap(void* h, int bv, int shapeA, const double* BA, const double* oA, int shapeB, const double* BB,
                   const double* oB) {
  System* s = (System*)h;
  Shape const& sa = s->shapes[shapeA];
  Shape const& sb = s->shapes[shapeB];
  int ca = 0, cb = 0;
  ConvexRef A{sa.type, sa.radius, sa.extent, &sa, &ca}, B{sb.type, sb.radius, sb.extent, &sb, &cb};
  Tf ta, tb;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      ta.B[i][j] = BA[3 * i + j];
      tb.B[i][j] = BB[3 * i + j];
    }
  ta.o = V3(oA[0], oA[1], oA[2]);
  tb.o = V3(oB[0], oB[1], oB[2]);
  return bvOverlap(bv, A, ta, B, tb) ? 1 : 0;
}

// <CollisionDetectionAlgorithm>: sv 1 = GJK_SV, tol = Tolerance,
// accel 1 = Acceleration ON, bv = BoundingVolume (0 none, 1 OBB, 2 OBC)
int orc_set_collision(void* h, int sv, double tol, int accel, int bv) {
  System* s = (System*)h;
  if (tol < 1e-10 || bv < 0 || bv > 2) return 1;
  s->col.sv = sv ? 1 : 0;
  s->col.tol = tol;
  s->col.accel = accel ? 1 : 0;
  s->col.bv = bv;
  return 0;
}

int orc_init(void* h) {
  System* s = (System*)h;
  g_col = s->col;
  double rmax = 0.;
  for (auto& p : s->pool) rmax = std::max(rmax, p.radius);
  if (rmax < 1.e-12) rmax = 1.e16;
  int e = buildGrid(s, 2. * (s->cellsize_factor * rmax));
  if (e) return s->err = e;
  // obstacle <-> cell links (LinkedCell.cpp:1371-1449, AABB test only)
  for (size_t io = 0; io < s->obstacles.size(); ++io) {
    Obs& o = s->obstacles[io];
    for (int m = 0; m < s->nb; ++m) {
      Cell& c = s->cells[m];
      bool in = std::fabs(o.aabb_c[0] - c.centre[0]) <= o.aabb_e[0] + s->edge[0] &&
                std::fabs(o.aabb_c[1] - c.centre[1]) <= o.aabb_e[1] + s->edge[1] &&
                std::fabs(o.aabb_c[2] - c.centre[2]) <= o.aabb_e[2] + s->edge[2];
      if (in && std::find(c.obstacles.begin(), c.obstacles.end(), (int)io) == c.obstacles.end())
        c.obstacles.push_back((int)io);
    }
  }
  bool per = s->periodic[0] || s->periodic[1] || s->periodic[2];
  std::vector<V3> vecs;
  size_t nmasters = s->pool.size();
  // clones are appended to the pool; at most 7 per master, plus the ones made
  // on later steps (never more than 7 alive per master), so reserve ahead so
  // that references into the pool stay valid
  s->pool.reserve(64 * nmasters + 64);
  for (size_t ip = 0; ip < nmasters; ++ip) {
    e = linkParticle(s, (int)ip);
    if (e) return s->err = e;
    if (per) {  // insertParticleSerial clone creation (LinkedCell.cpp:1826-1852)
      periodicVectors(s, s->cells[s->pool[ip].cell], vecs);
      for (auto& vec : vecs) {
        int ic = makeClone(s, (int)ip, vec);
        e = linkParticle(s, ic);
        if (e) return s->err = e;
        s->clones.insert({s->pool[ip].id, ic});
        s->active.push_back(ic);
      }
    }
    s->active.push_back((int)ip);  // AllComponents::WaitToActive
  }
  s->initialized = true;
  for (int ip : s->active) {
    Part& p = s->pool[ip];
    p.refp = p.x; p.F = p.weight; p.T = V3();
  }
  if (s->scheme == LEAPFROG) {
    e = computeParticlesForceAndTorque(s);
    if (e) return s->err = e;
  }
  return OK;
}

// Runs n steps; returns 0 or the first error code (ContactError 1,
// MotionError 2, SimulationError 3)
int orc_step(void* h, int n) {
  System* s = (System*)h;
  g_col = s->col;
  for (int i = 0; i < n; ++i) {
    int e = oneStep(s);
    if (e) return s->err = e;
  }
  return OK;
}

int orc_num_particles(void* h) { return (int)((System*)h)->insertion.size(); }
int orc_num_active(void* h) { return (int)((System*)h)->active.size(); }
double orc_time(void* h) { return ((System*)h)->time; }

// State of the masters, in id order.  Arrays hold 3 (or 4 for q: x, y, z, w)
// doubles per particle; null pointers are skipped.
void orc_get_state(void* h, double* x, double* v, double* w, double* wb, double* q, double* F, double* T) {
  System* s = (System*)h;
  int n = (int)s->insertion.size();
  for (int id = 1; id <= n; ++id) {
    Part& p = s->pool[s->master_of_id[id]];
    int i = id - 1;
    for (int a = 0; a < 3; ++a) {
      if (x) x[3 * i + a] = p.x[a];
      if (v) v[3 * i + a] = p.v[a];
      if (w) w[3 * i + a] = p.w[a];
      if (wb) wb[3 * i + a] = p.wb[a];
      if (F) F[3 * i + a] = p.F[a];
      if (T) T[3 * i + a] = p.T[a];
      if (q) q[4 * i + a] = p.q.v[a];
    }
    if (q) q[4 * i + 3] = p.q.w;
  }
}

// Per particle, the sum of the magnitudes of the contact forces of the last
// force evaluation (the normaliser of force-error figures).
void orc_get_force_sums(void* h, double* out) {
  System* s = (System*)h;
  int n = (int)s->insertion.size();
  for (int id = 1; id <= n; ++id) out[id - 1] = s->pool[s->master_of_id[id]].fsum;
}

// Per particle, the sum of the term scales of its memoryless contact forces
// in the last force evaluation (kn |pen| + (|gamman| + |gammat|) |v_rel|).
void orc_get_term_scales(void* h, double* out) {
  System* s = (System*)h;
  int n = (int)s->insertion.size();
  for (int id = 1; id <= n; ++id) out[id - 1] = s->pool[s->master_of_id[id]].tsum;
}

void orc_get_mass(void* h, double* mass, double* inertia, double* weight) {
  System* s = (System*)h;
  int n = (int)s->insertion.size();
  for (int id = 1; id <= n; ++id) {
    Part& p = s->pool[s->master_of_id[id]];
    if (mass) mass[id - 1] = p.mass;
    if (inertia) inertia[id - 1] = p.inertia[0];
    if (weight)
      for (int a = 0; a < 3; ++a) weight[3 * (id - 1) + a] = p.weight[a];
  }
}

void orc_get_obstacle_loads(void* h, double* F, double* T) {
  System* s = (System*)h;
  for (size_t i = 0; i < s->obstacles.size(); ++i)
    for (int a = 0; a < 3; ++a) {
      F[3 * i + a] = s->obstacles[i].F[a];
      T[3 * i + a] = s->obstacles[i].T[a];
    }
}

// Directed contact records of the last force evaluation: (receiver id,
// partner id); obstacles have negative ids.  Returns the count.
int orc_get_pairs(void* h, int* buf, int cap) {
  System* s = (System*)h;
  int n = (int)s->pairs.size();
  for (int i = 0; i < n && i < cap; ++i) {
    buf[2 * i] = s->pairs[i].self;
    buf[2 * i + 1] = s->pairs[i].other;
  }
  return n;
}

void orc_get_grid(void* h, int* dims, double* lc_origin, double* edge) {
  System* s = (System*)h;
  dims[0] = s->nbi; dims[1] = s->nbj; dims[2] = s->nbk;
  for (int a = 0; a < 3; ++a) { lc_origin[a] = s->lc_origin[a]; edge[a] = s->edge[a]; }
}

// Cell ijk of each master (Cell::GetCell on the current positions)
void orc_get_cells(void* h, int* ijk) {
  System* s = (System*)h;
  int n = (int)s->insertion.size();
  for (int id = 1; id <= n; ++id) {
    int c[3];
    GetCell(s, s->pool[s->master_of_id[id]].x, c);
    for (int a = 0; a < 3; ++a) ijk[3 * (id - 1) + a] = c[a];
  }
}

void orc_get_stats(void* h, double* out) {
  System* s = (System*)h;
  out[0] = s->overlap_max;
  out[1] = s->overlap_mean;
  out[2] = (double)s->n_recovered;
  out[3] = (double)s->err_id0;
  out[4] = (double)s->err_id1;
  out[5] = s->time_overlap_max;
  out[6] = s->gjk_iter_mean;
  out[7] = s->overlap_count;
  out[8] = s->gjk_count;
}

// Overwrites master states (id order) and re-links, as a reload would; clones
// are rebuilt.  Forces are recomputed if recompute != 0.
int orc_set_state(void* h, const double* x, const double* v, const double* w, const double* q, int recompute) {
  System* s = (System*)h;
  g_col = s->col;
  int n = (int)s->insertion.size();
  // drop clones and all cell links
  for (auto& kv : s->clones) s->pool[kv.second].alive = false;
  s->clones.clear();
  std::list<int> masters;
  for (int id = 1; id <= n; ++id) masters.push_back(s->master_of_id[id]);
  for (auto& c : s->cells) c.parts.clear();
  s->active.clear();
  bool per = s->periodic[0] || s->periodic[1] || s->periodic[2];
  std::vector<V3> vecs;
  for (int ip : masters) {
    Part& p = s->pool[ip];
    int i = p.id - 1;
    p.clone = false;
    if (x) p.x = V3(x[3 * i], x[3 * i + 1], x[3 * i + 2]);
    if (q) p.q = Q(V3(q[4 * i], q[4 * i + 1], q[4 * i + 2]), q[4 * i + 3]);
    if (v) p.v = V3(v[3 * i], v[3 * i + 1], v[3 * i + 2]);
    if (w) {
      p.w = V3(w[3 * i], w[3 * i + 1], w[3 * i + 2]);
      p.wb = multToVector3(conj(p.q), vecq(p.w, p.q));
    }
    int e = linkParticle(s, ip);
    if (e) return s->err = e;
    if (per) {
      periodicVectors(s, s->cells[p.cell], vecs);
      for (auto& vec : vecs) {
        int ic = makeClone(s, ip, vec);
        e = linkParticle(s, ic);
        if (e) return s->err = e;
        s->clones.insert({s->pool[ip].id, ic});
        s->active.push_back(ic);
      }
    }
    s->active.push_back(ip);
  }
  if (recompute) {
    int e = computeParticlesForceAndTorque(s);
    if (e) return s->err = e;
  }
  return OK;
}

// Overwrites F and T (id order), e.g. to integrate with prescribed loads.
void orc_set_forces(void* h, const double* F, const double* T) {
  System* s = (System*)h;
  int n = (int)s->insertion.size();
  for (int id = 1; id <= n; ++id) {
    Part& p = s->pool[s->master_of_id[id]];
    int i = id - 1;
    p.F = V3(F[3 * i], F[3 * i + 1], F[3 * i + 2]);
    p.T = V3(T[3 * i], T[3 * i + 1], T[3 * i + 2]);
  }
}

// Kick-drift and kick stages on their own (integrator tests with prescribed
// F, T): ParticleKinematics::Move and ::advanceVelocity over all masters.
int orc_move_only(void* h, double dtv, double dtx) {
  System* s = (System*)h;
  s->step_counter++;
  for (int ip : s->active) {
    Part& p = s->pool[ip];
    if (p.clone) continue;
    double depl = kinMove(s, p, dtv, dtx);
    if (depl > p.crust) { s->err_id0 = p.id; return s->err = MOTION_ERROR; }
  }
  return OK;
}
void orc_advance_only(void* h, double dtv) {
  System* s = (System*)h;
  for (int ip : s->active) {
    Part& p = s->pool[ip];
    if (!p.clone) kinAdvance(s, p, dtv);
  }
}

const char* orc_message(void* h) { return ((System*)h)->msg; }

}  // extern "C"
