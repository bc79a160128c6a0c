// The C interface through which Basilisk's DLMFD solver drives Grains3D
// (GRAINS_V1/Grains/InterfaceCBasilisk/include/InterfaceGrains3DBasilisk.h:
// 12-32, called from Octree/DLMFD/DLMFD_Grains3D.h:34-169), implemented over
// the GPU system.  One static system per process, as in the reference
// (InterfaceGrains3DBasilisk.cpp:15); the calls are not re-entrant.
// Configuration errors end the process with status 1 (Grains::grainsAbort);
// a ContactError, MotionError or SimulationError inside Simu_Grains ends the
// sub-cycling with a message (GrainsCoupledWithFluid.cpp:201-230).
#include <climits>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/InterfaceGrains3DBasilisk.h"
#include "../../include/grains_gpu.h"
#include "dem_system.hpp"
#include "grains_xml.hpp"

namespace {

grains_gpu_system* g_sys = nullptr;
int g_cycle0 = 0;
bool g_error = false;

void die(const char* what) {
  std::fprintf(stderr, "Grains: %s\n", what);
  std::exit(1);
}

grains_gpu_system* sys() {
  if (!g_sys) die("Init_Grains has not been called");
  return g_sys;
}

// GrainsExec::doubleToString(ios::scientific, FORMAT16DIGITS, x)
// (GrainsExec.cpp:212-224): 16 digits after the point, plain "0" for zero
std::string sci16(double x) {
  if (x == 0.) return std::signbit(x) ? "-0" : "0";
  char b[40];
  std::snprintf(b, sizeof(b), "%.16e", x);
  return b;
}
// default ostream formatting of a double (6 significant digits)
std::string plain(double x) {
  std::ostringstream o;
  o << x;
  return o.str();
}

// Matrix product in the reference's order (Matrix.cpp:537-550)
void matmul(const double* a, const double* b, double* c) {
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j)
      c[3 * i + j] = a[3 * i] * b[j] + a[3 * i + 1] * b[3 + j] + a[3 * i + 2] * b[6 + j];
}

// Box corner numbering and faces of Box::setCornersFaces (Box.cpp:92-160)
const int kBoxFaces[6][4] = {{4, 5, 6, 7}, {5, 1, 2, 6}, {1, 0, 3, 2}, {0, 4, 7, 3}, {4, 0, 1, 5}, {3, 7, 6, 2}};
void box_corners(const double* e, double* out /*8x3*/) {
  for (int n = 0; n < 8; ++n) {
    const bool px = n == 1 || n == 2 || n == 5 || n == 6, py = n >= 4, pz = n == 2 || n == 3 || n == 6 || n == 7;
    out[3 * n] = px ? e[0] : -e[0];
    out[3 * n + 1] = py ? e[1] : -e[1];
    out[3 * n + 2] = pz ? e[2] : -e[2];
  }
}

// Transform::operator() (Transform.cpp:312-317)
void apply(const double* B, const double* o, const double* p, double* q) {
  for (int i = 0; i < 3; ++i) q[i] = B[3 * i] * p[0] + B[3 * i + 1] * p[1] + B[3 * i + 2] * p[2] + o[i];
}

// RigidBody::writePositionInFluid (RigidBody.cpp:445-500), 3D
void write_envelope(std::ostringstream& o, const double* B, const double* x, const std::vector<double>& pts,
                    const std::vector<std::vector<int>>& faces) {
  const size_t np = pts.size() / 3;
  o << " " << np << "\n";
  for (size_t k = 0; k < np; ++k) {
    double q[3];
    apply(B, x, &pts[3 * k], q);
    o << sci16(q[0]) << " " << sci16(q[1]) << " " << sci16(q[2]) << "\n";
  }
  if (faces.empty()) {
    o << "0\n";
  } else {
    o << faces.size() << "\n";
    for (auto& f : faces) {
      o << f.size() << " ";
      for (int i : f) o << i << " ";
      o << "\n";
    }
  }
}

void write_kinematics(std::ostringstream& o, const double* v, const double* w) {
  for (int a = 0; a < 3; ++a) o << sci16(v[a]) << " ";
  for (int a = 0; a < 3; ++a) o << sci16(w[a]) << " ";
}

// GrainsCoupledWithFluid::GrainsToFluid (GrainsCoupledWithFluid.cpp:912-1120)
int to_fluid_text(grains_gpu_system* s, std::string& text) {
  const size_t n = (size_t)s->nparticles;
  std::vector<double> x(3 * n), v(3 * n), w(3 * n), B(9 * n);
  int e = grains_gpu_get_state(s, x.data(), v.data(), w.data(), nullptr, nullptr, nullptr, nullptr);
  if (!e) e = grains_gpu_get_basis(s, B.data());
  if (e) return e;
  size_t nobs = 0;
  for (int f : s->obs_to_fluid) nobs += f ? 1 : 0;
  std::ostringstream o;
  o << n + nobs << "\n";
  const dem::Grid& G = s->grid;
  for (size_t i = 0; i < n; ++i) {
    const int sh = s->h_shape[i];
    const int type = sh < 0 ? dem::kShapeSphere : s->shapes[sh].type;
    // corners, envelope and faces of the convex (Sphere.cpp:116-145,
    // Box.cpp:226-279, Polytope.cpp:92-95)
    std::vector<double> pts;
    std::vector<std::vector<int>> faces;
    int ncorners = 1;
    if (type == dem::kShapeSphere) {
      pts = {0., 0., 0.};
    } else if (type == dem::kShapeBox) {
      pts.resize(24);
      box_corners(s->shapes[sh].extent, pts.data());
      for (auto& f : kBoxFaces) faces.push_back(std::vector<int>(f, f + 4));
      ncorners = 666;
    } else {
      const dem::ShapeDesc& d = s->shapes[sh];
      pts.assign(s->shape_verts.begin() + 3 * (size_t)d.vert_off,
                 s->shape_verts.begin() + 3 * (size_t)(d.vert_off + d.nv));
      const std::vector<int>& fl = s->shape_faces[sh];
      for (size_t k = 0; k < fl.size();) {
        int m = fl[k++];
        faces.push_back(std::vector<int>(fl.begin() + k, fl.begin() + k + m));
        k += m;
      }
      ncorners = d.nv;
    }
    // Particle::computeInertiaTensorSpaceFixed (Particle.cpp:1797-1821)
    double Ib[9];
    if (sh < 0) {
      const double I = s->h_inertia[i];
      const double t[9] = {I, 0., 0., 0., I, 0., 0., 0., I};
      std::memcpy(Ib, t, sizeof(Ib));
    } else {
      const double* u = &s->shape_inertia[6 * (size_t)sh];
      const double r = s->h_density[i];
      const double I6[6] = {u[0] * r, u[1] * r, u[2] * r, u[3] * r, u[4] * r, u[5] * r};
      const double t[9] = {I6[0], I6[1], I6[2], I6[1], I6[3], I6[4], I6[2], I6[4], I6[5]};
      std::memcpy(Ib, t, sizeof(Ib));
    }
    const double* mr = &B[9 * i];
    double mrt[9], tmp[9], Is[9];
    for (int a = 0; a < 3; ++a)
      for (int b = 0; b < 3; ++b) mrt[3 * a + b] = mr[3 * b + a];
    matmul(Ib, mrt, tmp);
    matmul(mr, tmp, Is);
    // periodic clones of a particle in a buffer cell (App.cpp:65-240)
    std::vector<std::vector<double>> vecs;
    {
      int axes[3], side[3], na = 0;
      for (int a = 0; a < 3; ++a) {
        if (!G.periodic[a]) continue;
        int ri = int(std::floor((x[3 * i + a] - G.lo[a]) / G.edge[a]));
        int sd = ri == 1 ? -1 : (ri == G.n[a] ? 1 : 0);
        if (sd) {
          axes[na] = a;
          side[na++] = sd;
        }
      }
      for (int mask = 1; mask < (1 << na); ++mask) {
        double vec[3] = {0., 0., 0.};
        for (int b = 0; b < na; ++b)
          if (mask & (1 << b)) vec[axes[b]] = side[b] < 0 ? G.L[axes[b]] : -G.L[axes[b]];
        std::vector<double> pv(3);
        for (int a = 0; a < 3; ++a) pv[a] = (x[3 * i + a] + vec[a]) - x[3 * i + a];
        vecs.push_back(pv);
      }
    }
    o << i << " " << ncorners << "\n";
    o << (vecs.empty() ? "P" : "PP") << " ";
    write_kinematics(o, &v[3 * i], &w[3 * i]);
    o << sci16(s->h_density[i]) << " " << sci16(s->h_mass[i]) << " ";
    const int I6idx[6] = {0, 1, 2, 4, 5, 8};
    for (int k : I6idx) o << sci16(Is[k]) << " ";
    for (int k = 0; k < 9; ++k) o << sci16(mr[k]) << " ";
    o << sci16(x[3 * i]) << " " << sci16(x[3 * i + 1]) << " " << sci16(x[3 * i + 2]) << "\n";
    if (!vecs.empty()) {
      o << vecs.size() << "\n";
      for (auto& pv : vecs) o << sci16(pv[0]) << " " << sci16(pv[1]) << " " << sci16(pv[2]) << "\n";
    }
    o << plain(s->h_rad[i]);
    write_envelope(o, mr, &x[3 * i], pts, faces);
  }
  size_t cid = n;
  for (size_t k = 0; k < s->obstacles.size(); ++k) {
    if (!s->obs_to_fluid[k]) continue;
    const dem::BoxObstacle& ob = s->obstacles[k];
    const double zero[3] = {0., 0., 0.};
    o << cid++ << " " << 666 << "\n";
    o << "O ";
    write_kinematics(o, zero, zero);
    o << "1000. 0. 0. 0. 0. 0. 0. 0. ";
    for (int j = 0; j < 9; ++j) o << sci16(ob.B[j]) << " ";
    o << sci16(ob.centre[0]) << " " << sci16(ob.centre[1]) << " " << sci16(ob.centre[2]) << "\n";
    const double* e3 = ob.extent;
    o << plain(std::sqrt(e3[0] * e3[0] + e3[1] * e3[1] + e3[2] * e3[2]));
    std::vector<double> pts(24);
    box_corners(e3, pts.data());
    std::vector<std::vector<int>> faces;
    for (auto& f : kBoxFaces) faces.push_back(std::vector<int>(f, f + 4));
    write_envelope(o, ob.B, ob.centre, pts, faces);
  }
  text = o.str();
  return GRAINS_OK;
}

}  // namespace

extern "C" {

int grains_gpu_create_coupled(const char* xml_path, double fluid_rho, int fluid_corrected, grains_gpu_system** out) {
  if (!xml_path || !out) return GRAINS_ARGUMENT_ERROR;
  dem::XmlOptions opt;
  opt.coupled = true;
  opt.fluid_density = fluid_rho;
  opt.track_basis = true;
  opt.fluid_corrected = fluid_corrected != 0;
  return dem::create_from_xml(xml_path, GRAINS_FP64, opt, out);
}

// GrainsCoupledWithFluid::Simulation (GrainsCoupledWithFluid.cpp:123-232):
// dt_eff = max(min(maxdt, dt_fluid / ndt), mindt), ndt = int(dt_fluid / dt_eff),
// dt = dt_fluid / ndt, then ndt kick-drift-kick steps
int grains_gpu_fluid_step(grains_gpu_system* s, double dt_fluid, int* nsub) {
  if (!s || !s->initialized || !(dt_fluid > 0.) || s->coupled_ndt == 0) return GRAINS_ARGUMENT_ERROR;
  double dteff = std::max(std::min(s->coupled_max_dt, dt_fluid / double(s->coupled_ndt)), s->coupled_min_dt);
  s->coupled_ndt = size_t(dt_fluid / dteff);
  if (nsub) *nsub = (int)s->coupled_ndt;
  if (s->coupled_ndt == 0) return GRAINS_OK;
  s->cfg.dt = dt_fluid / double(s->coupled_ndt);
  return grains_gpu_step(s, (int)s->coupled_ndt);
}

int grains_gpu_to_fluid(grains_gpu_system* s, char** buf, int* size) {
  if (!s || !s->initialized || !buf || !size) return GRAINS_ARGUMENT_ERROR;
  if (!s->has_shapes()) {
    s->last_error = "the system does not track particle bases (build it with grains_gpu_create_coupled)";
    return GRAINS_ARGUMENT_ERROR;
  }
  std::string text;
  int e = to_fluid_text(s, text);
  if (e) return e;
  // every token is followed by one space: at most text.size() + 1 bytes
  // before the terminating NUL
  const size_t cap = text.size() + 2;
  char* out = (char*)std::malloc(cap);
  if (!out) return GRAINS_DEVICE_ERROR;
  std::istringstream iss(text);
  std::string tok;
  size_t pos = 0;
  while (iss >> tok && pos + tok.size() + 1 < cap) {
    std::memcpy(out + pos, tok.data(), tok.size());
    pos += tok.size();
    out[pos++] = ' ';
  }
  out[pos] = '\0';
  *size = int(pos) + 1;
  *buf = out;
  return GRAINS_OK;
}

// GrainsCoupledWithFluid::updateParticlesVelocity (GrainsCoupledWithFluid.cpp:
// 1126-1181): rows of (vx, vy, vz, wx, wy, wz) by particle id - 1
int grains_gpu_set_velocities(grains_gpu_system* s, const double* vw, size_t m) {
  if (!s || !s->initialized || !vw) return GRAINS_ARGUMENT_ERROR;
  const size_t n = (size_t)s->nparticles;
  std::vector<double> v(3 * n), w(3 * n);
  int e = grains_gpu_get_state(s, nullptr, v.data(), w.data(), nullptr, nullptr, nullptr, nullptr);
  if (e) return e;
  for (size_t i = 0; i < n && i < m; ++i)
    for (int a = 0; a < 3; ++a) {
      v[3 * i + a] = vw[6 * i + a];
      w[3 * i + a] = vw[6 * i + 3 + a];
    }
  return grains_gpu_set_state(s, nullptr, v.data(), w.data(), nullptr, 0);
}

int grains_gpu_count_rigid_bodies(grains_gpu_system* s, size_t* nparticles, size_t* nobstacles) {
  if (!s) return GRAINS_ARGUMENT_ERROR;
  if (nparticles) *nparticles = (size_t)s->nparticles;
  if (nobstacles) {
    size_t k = 0;
    for (int f : s->obs_to_fluid) k += f ? 1 : 0;
    *nobstacles = k;
  }
  return GRAINS_OK;
}

// ---- the Basilisk entry points over one static system ----

void Init_Grains(char const* inputfile, double fluid_rho, const bool b_restart, const bool b_fluidcorrectedacc) {
  if (g_sys) die("Init_Grains called twice");
  grains_gpu_system* s = nullptr;
  int e = grains_gpu_create_coupled(inputfile, fluid_rho, b_fluidcorrectedacc ? 1 : 0, &s);
  if (e) {
    std::string msg = s ? s->last_error : std::string("cannot build the system");
    die(msg.c_str());
  }
  // GrainsCoupledWithFluid::setReloadSame: the last reload file of the table
  if (b_restart && grains_gpu_reload_same(s)) die(grains_gpu_last_error(s));
  g_sys = s;
  std::printf("Construction of Grains completed\n");
}

void Simu_Grains(const double dt_fluid) {
  grains_gpu_system* s = sys();
  int e = grains_gpu_fluid_step(s, dt_fluid, nullptr);
  if (e) {
    std::fprintf(stderr, "Grains: %s\n", grains_gpu_last_error(s));
    g_error = true;
  }
}

char* GrainsToBasilisk(int* pstrsize) {
  char* buf = nullptr;
  if (grains_gpu_to_fluid(sys(), &buf, pstrsize)) die(grains_gpu_last_error(sys()));
  return buf;
}

void SetInitialTime(double tinit) { grains_gpu_set_time(sys(), tinit); }

// InterfaceGrains3DBasilisk.cpp:83-92: the first call starts the
// post-processing writers (InitialPostProcessing -> PostProcessing_start,
// files truncated), the next ones append a frame (doPostProcessing).  The
// writer this library has is RawData (<PostProcessing><Writers><RawData>
// of the input file): one frame of the current state, coordination
// numbers included.
void SaveResults_Grains(void) {
  static unsigned int ppcounter = 0;
  grains_gpu_system* s = sys();
  if (grains_gpu_post_processing(s, ppcounter == 0 ? 1 : 0)) die(grains_gpu_last_error(s));
  ++ppcounter;
}

// GrainsCoupledWithFluid::checkParaviewPostProcessing prepares the Paraview
// writer's directory; this library has no Paraview writer, so there is
// nothing to check.
void checkParaviewPostProcessing_Grains(char* solid_resDir) { (void)solid_resDir; }

void UpdateVelocityGrains(double** arrayv, const int m) {
  grains_gpu_system* s = sys();
  const size_t n = (size_t)s->nparticles;
  if ((size_t)m != n)
    std::printf("WARNING: numbers of particles in Grains and in the fluid solver are different in "
                "UpdateVelocityGrains\n");
  std::vector<double> vw(6 * (size_t)m);
  for (int i = 0; i < m; ++i)
    for (int j = 0; j < 6; ++j) vw[6 * (size_t)i + j] = arrayv[i][j];
  if (grains_gpu_set_velocities(s, vw.data(), (size_t)m)) die(grains_gpu_last_error(s));
}

void SetInitialCycleNumber(int cycle0) {
  sys();
  g_cycle0 = cycle0;
}

void NumberOfRigidBodiesInBasilisk(size_t* nparticles, size_t* nobstacles) {
  grains_gpu_count_rigid_bodies(sys(), nparticles, nobstacles);
}

}  // extern "C"
