// Particle insertion (Grains::insertParticle, Grains.cpp:2030-2140): see
// grains_insert.hpp.  The random numbers come from the C library's
// random()/rand() stream seeded by <RandomGeneratorSeed>, drawn in the
// reference's order: class pick, position, rotation, velocities, then the
// overlap test.
#include "grains_insert.hpp"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <string>

#include "../../include/grains_gpu.h"
#include "dem_device.hpp"
#include "dem_system.hpp"

using namespace dem;

namespace grains_insert {
namespace {

constexpr double kPi = 3.1415926535897932384626433832795;

double unit_random() { return double(random()) / double(INT_MAX); }

// Window::getInsertionPoint (Window.cpp:238-320)
void window_point(const Window& w, double* P) {
  if (w.type == kBox) {
    for (int a = 0; a < 3; ++a) P[a] = w.A[a] + unit_random() * (w.B[a] - w.A[a]);
  } else if (w.type == kLine) {
    const double t = unit_random();
    for (int a = 0; a < 3; ++a) P[a] = w.A[a] + t * (w.B[a] - w.A[a]);
  } else {
    const double r = w.type == kCylinder ? unit_random() * w.radius
                                         : w.radius_int + unit_random() * (w.radius - w.radius_int);
    const double theta = unit_random() * 2. * kPi;
    const double axial = unit_random() * w.height;
    if (w.axis == 0) {
      P[0] = w.A[0] + axial;
      P[1] = w.A[1] + r * std::cos(theta);
      P[2] = w.A[2] + r * std::sin(theta);
    } else if (w.axis == 1) {
      P[0] = w.A[0] + r * std::sin(theta);
      P[1] = w.A[1] + axial;
      P[2] = w.A[2] + r * std::cos(theta);
    } else {
      P[0] = w.A[0] + r * std::cos(theta);
      P[1] = w.A[1] + r * std::sin(theta);
      P[2] = w.A[2] + axial;
    }
  }
}

// m * n, each entry summed over k in order (Matrix::operator*)
void mat_mul(const double* m, const double* n, double* out) {
  double t[9];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) t[3 * i + j] = m[3 * i] * n[j] + m[3 * i + 1] * n[3 + j] + m[3 * i + 2] * n[6 + j];
  for (int k = 0; k < 9; ++k) out[k] = t[k];
}

// GrainsExec::RandomRotationMatrix (GrainsExec.cpp:476-502), three dimensions
void random_rotation(double* R) {
  const double az = 2. * kPi * (double)rand() / RAND_MAX;
  const double rZ[9] = {std::cos(az), -std::sin(az), 0., std::sin(az), std::cos(az), 0., 0., 0., 1.};
  const double ax = 2. * kPi * (double)rand() / RAND_MAX;
  const double rX[9] = {1., 0., 0., 0., std::cos(ax), -std::sin(ax), 0., std::sin(ax), std::cos(ax)};
  const double ay = 2. * kPi * (double)rand() / RAND_MAX;
  const double rY[9] = {std::cos(ay), 0., std::sin(ay), 0., 1., 0., -std::sin(ay), 0., std::cos(ay)};
  double tmp[9];
  mat_mul(rY, rZ, tmp);
  mat_mul(rX, tmp, R);
}

// the geometry of a body of shape class `shape` (-1: a sphere of radius rad)
ConvexGeo body_geo(const grains_gpu_system* s, int shape, double rad, double crust, const double* x,
                   const double* B) {
  ConvexGeo g{};
  g.o = d3{x[0], x[1], x[2]};
  for (int j = 0; j < 9; ++j) g.B[j] = B ? B[j] : ((j % 4) == 0 ? 1. : 0.);
  g.opt = s->gjk_opts;
  if (shape < 0) {
    g.type = kShapeSphere;
    g.r = g.circR = rad;
    g.crust = crust;
    g.scaling = d3{1., 1., 1.};
    return g;
  }
  const ShapeDesc& d = s->shapes[(size_t)shape];
  g.type = d.type;
  g.nv = d.nv;
  g.verts = s->shape_verts.data() + 3 * (size_t)d.vert_off;
  g.r = d.radius;
  g.e = d3{d.extent[0], d.extent[1], d.extent[2]};
  g.scaling = d3{d.scaling[0], d.scaling[1], d.scaling[2]};
  g.bext = d3{d.bext[0], d.bext[1], d.bext[2]};
  g.crust = d.crust;
  g.circR = d.circR;
  g.obb = d3{d.obb[0], d.obb[1], d.obb[2]};
  return g;
}

// a box obstacle as a convex (the device's convex_of_obstacle)
ConvexGeo obstacle_geo(const grains_gpu_system* s, const BoxObstacle& o) {
  ConvexGeo g{};
  g.type = kShapeBox;
  g.e = d3{o.extent[0], o.extent[1], o.extent[2]};
  for (int j = 0; j < 9; ++j) g.B[j] = o.B[j];
  g.o = d3{o.centre[0], o.centre[1], o.centre[2]};
  const double* e = o.extent;
  g.scaling = d3{e[0] < kEps ? 1. : (e[0] - o.crust) / e[0], e[1] < kEps ? 1. : (e[1] - o.crust) / e[1],
                 e[2] < kEps ? 1. : (e[2] - o.crust) / e[2]};
  g.bext = g.e;
  g.crust = o.crust;
  g.circR = norm(g.e);
  g.obb = g.e;
  g.opt = s->gjk_opts;
  return g;
}

// RigidBodyWithCrust::isContact (RigidBodyWithCrust.cpp:842-955) of the
// component A already in place with the new body B: sphere pairs by their
// radii, sphere-box pairs by Box::IntersectionPointSPHERE (a centre inside
// the box counts as contact), other pairs by the circumscribed spheres (and
// the bounding volume if one is set), then Johnson's GJK on the crust-shrunk
// bodies against the crusts' sum less EPSILON; bv_only stops after the
// bounding volumes.
bool touching(const ConvexGeo& A, const ConvexGeo& B, bool bv_only) {
  const bool sa = A.type == kShapeSphere, sb = B.type == kShapeSphere;
  if (sa && sb) return norm(B.o - A.o) - (A.circR + B.circR) < 0.;
  if ((sa && B.type == kShapeBox) || (A.type == kShapeBox && sb)) {
    const ConvexGeo& bx = sa ? B : A;
    const ConvexGeo& sp = sa ? A : B;
    double Binv[9], io[3];
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) Binv[3 * i + j] = bx.B[3 * j + i];
    for (int i = 0; i < 3; ++i) io[i] = dot(-d3{Binv[3 * i], Binv[3 * i + 1], Binv[3 * i + 2]}, bx.o);
    const double e[3] = {bx.e.x, bx.e.y, bx.e.z};
    double overlap = 0.;
    d3 cp;
    if (!box_sphere(e, apply_tf(Binv, io, sp.o), sp.circR, overlap, cp)) return true;
    return overlap < 0.;
  }
  bool bv = norm(A.o - B.o) < A.circR + B.circR;
  if (A.opt.bv && bv) bv = bv_overlap(A, B);
  if (bv_only || !bv) return bv;
  const Tf a2w = tf_scaled(A, A.scaling), b2w = tf_scaled(B, B.scaling);
  d3 pa, pb;
  int it = 0;
  const double d = gjk_closest(A, a2w, B, b2w, pa, pb, it, -1., nullptr, true);
  return d < A.crust + B.crust - kEps;
}

// cell coordinates of a point (wrapped on periodic axes) and their key
void cell_of(const grains_gpu_system* s, const Plan& p, const double* x, long* k) {
  for (int a = 0; a < 3; ++a) {
    k[a] = (long)std::floor((x[a] - s->cfg.origin[a]) / p.h[a]);
    if (p.ncell[a] > 0) k[a] = ((k[a] % p.ncell[a]) + p.ncell[a]) % p.ncell[a];
  }
}
long long cell_key(const long* k) {
  return ((long long)(k[0] & 0x1FFFFF) << 42) | ((long long)(k[1] & 0x1FFFFF) << 21) | (long long)(k[2] & 0x1FFFFF);
}

// the cell edge: the largest circumscribed diameter among the classes and
// the particles in place
void setup_cells(const grains_gpu_system* s, Plan& p) {
  double d = 0.;
  for (const Class& c : p.classes) d = std::max(d, 2. * (c.shape < 0 ? c.radius : s->shapes[(size_t)c.shape].circR));
  for (size_t j = 0; j < (size_t)s->nparticles; ++j) {
    const int shape = s->h_shape.empty() ? -1 : s->h_shape[j];
    d = std::max(d, 2. * (shape < 0 ? s->h_rad[j] : s->shapes[(size_t)shape].circR));
  }
  for (int a = 0; a < 3; ++a) {
    const double L = s->cfg.max[a] - s->cfg.origin[a];
    p.h[a] = d > 0. ? d : 1.;
    p.ncell[a] = 0;
    if (s->cfg.periodic[a]) {
      p.ncell[a] = std::max(1L, (long)std::floor(L / p.h[a]));
      p.h[a] = L / double(p.ncell[a]);
    }
  }
  p.cells.clear();
  for (size_t j = 0; j < (size_t)s->nparticles; ++j) {
    long k[3];
    cell_of(s, p, &s->h_x[3 * j], k);
    p.cells[cell_key(k)].push_back((int)j);
  }
}

// whether the new body touches particle j (its nearest periodic image)
bool touches_particle(const grains_gpu_system* s, const Plan& p, const ConvexGeo& nb, const std::vector<double>& x,
                      const std::vector<double>& B, size_t j) {
  double L[3];
  for (int a = 0; a < 3; ++a) L[a] = s->cfg.periodic[a] ? s->cfg.max[a] - s->cfg.origin[a] : 0.;
  const double c[3] = {nb.o.x, nb.o.y, nb.o.z};
  const int shape = s->h_shape.empty() ? -1 : s->h_shape[j];
  const double reach = nb.circR + (shape < 0 ? s->h_rad[j] : s->shapes[(size_t)shape].circR);
  double xj[3];
  for (int a = 0; a < 3; ++a) {
    const double dx = c[a] - x[3 * j + a];
    double sh = 0.;
    if (L[a] > 0. && dx > 0.5 * L[a]) sh = L[a];
    else if (L[a] > 0. && dx < -0.5 * L[a]) sh = -L[a];
    xj[a] = x[3 * j + a] + sh;
    if (!(std::fabs(dx - sh) < reach)) return false;
  }
  const ConvexGeo A = body_geo(s, shape, s->h_rad[j], s->h_crust[j], xj, B.empty() ? nullptr : &B[9 * j]);
  return touching(A, nb, p.bv_only);
}

// whether the new body touches a particle in place or an obstacle: before
// initialisation through the cells, afterwards over all particles
bool blocked(const grains_gpu_system* s, const Plan& p, const ConvexGeo& nb, const std::vector<double>& x,
             const std::vector<double>& B) {
  if (!s->initialized) {
    const double c[3] = {nb.o.x, nb.o.y, nb.o.z};
    long k0[3];
    cell_of(s, p, c, k0);
    long long seen[27];
    int nseen = 0;
    for (int dx = -1; dx <= 1; ++dx)
      for (int dy = -1; dy <= 1; ++dy)
        for (int dz = -1; dz <= 1; ++dz) {
          long k[3] = {k0[0] + dx, k0[1] + dy, k0[2] + dz};
          for (int a = 0; a < 3; ++a)
            if (p.ncell[a] > 0) k[a] = ((k[a] % p.ncell[a]) + p.ncell[a]) % p.ncell[a];
          const long long key = cell_key(k);
          bool dup = false;
          for (int i = 0; i < nseen; ++i) dup |= seen[i] == key;
          if (dup) continue;
          seen[nseen++] = key;
          auto it = p.cells.find(key);
          if (it == p.cells.end()) continue;
          for (int j : it->second)
            if (touches_particle(s, p, nb, x, B, (size_t)j)) return true;
        }
  } else {
    for (size_t j = 0; j < x.size() / 3; ++j)
      if (touches_particle(s, p, nb, x, B, j)) return true;
  }
  for (const BoxObstacle& o : s->obstacles)
    if (touching(obstacle_geo(s, o), nb, p.bv_only)) return true;
  return false;
}

}  // namespace

int attempt(grains_gpu_system* s, Plan& p, bool* inserted) {
  *inserted = false;
  int e = GRAINS_OK;
  if (p.counter == 0 && p.pending() > 0) {
    // AllComponents::getParticleToInsert (AllComponents.cpp:434-480)
    if (p.wait < 0) {
      size_t type = 0;
      if (p.random_order) {
        std::vector<size_t> left;
        for (size_t i = 0; i < p.classes.size(); ++i)
          if (p.classes[i].remaining) left.push_back(i);
        const double v = unit_random();
        type = left[std::min(left.size() - 1, (size_t)(double(left.size()) * v))];
      } else {
        while (!p.classes[type].remaining) ++type;
      }
      p.wait = (int)type;
      for (int j = 0; j < 9; ++j) p.wait_B[j] = p.classes[type].B[j];
    }
    const Class& c = p.classes[(size_t)p.wait];
    // Grains::getInsertionPoint (Grains.cpp:2033-2064)
    double P[3];
    size_t isp = 0;
    const size_t npos = p.positions.size() / 3;
    if (npos) {
      isp = npos - 1;
      if (p.random_order) isp = std::min(npos - 1, (size_t)(unit_random() * double(npos)));
      for (int a = 0; a < 3; ++a) P[a] = p.positions[3 * isp + a];
    } else {
      size_t iw = 0;
      if (p.windows.size() != 1) {
        iw = (size_t)(unit_random() * double(p.windows.size()));
        if (iw == p.windows.size()) --iw;
      }
      window_point(p.windows[iw], P);
    }
    if (p.random_angle) {
      double R[9];
      random_rotation(R);
      mat_mul(p.wait_B, R, p.wait_B);
    } else if (p.angle_next < p.angles.size() / 9) {
      mat_mul(p.wait_B, &p.angles[9 * p.angle_next], p.wait_B);
    }
    // Grains::computeInitVit (Grains.cpp:2148-2200)
    double v[3] = {0., 0., 0.}, w[3] = {0., 0., 0.};
    if (p.vmode == 1) {
      for (int a = 0; a < 3; ++a) {
        v[a] = p.vconst[a];
        w[a] = p.wconst[a];
      }
    } else if (p.vmode == 2) {
      if (p.amp_t != 0.)
        for (int a = 0; a < 3; ++a) v[a] = p.amp_t * (2. * unit_random() - 1.);
      if (p.amp_r != 0.) {
        w[2] = p.amp_r * (2. * unit_random() - 1.);
        w[0] = p.amp_r * (2. * unit_random() - 1.);
        w[1] = p.amp_r * (2. * unit_random() - 1.);
      }
    }
    ++p.attempts;
    // the particles in place: the construction data before initialisation,
    // the current state afterwards
    std::vector<double> xs, Bs;
    const size_t n = (size_t)s->nparticles;
    if (s->initialized) {
      xs.resize(3 * n);
      if ((e = grains_gpu_get_state(s, xs.data(), nullptr, nullptr, nullptr, nullptr, nullptr, nullptr))) return e;
      if (s->has_shapes()) {
        Bs.resize(9 * n);
        if ((e = grains_gpu_get_basis(s, Bs.data()))) return e;
      }
    } else {
      xs = s->h_x;
      if (s->has_shapes()) Bs = s->h_B;
    }
    if (!s->initialized && p.h[0] == 0.) setup_cells(s, p);
    const ConvexGeo nb = body_geo(s, c.shape, c.radius, c.crust, P, p.wait_B);
    if (p.force || !blocked(s, p, nb, xs, Bs)) {
      s->next_class = c.xml_class;
      auto add = [&]() -> int {
        int mat = c.mat;
        if (c.shape < 0) return grains_gpu_add_spheres(s, 1, &c.radius, &c.density, &c.crust, &mat, P, v, w);
        return grains_gpu_add_particles(s, c.shape, 1, &c.density, &mat, P, v, w, p.wait_B);
      };
      const bool live = s->initialized;
      e = live ? s->grow(add) : add();
      if (e) return e;
      if (!live) {
        long k[3];
        cell_of(s, p, P, k);
        p.cells[cell_key(k)].push_back(s->nparticles - 1);
      }
      if (npos) {
        if (p.random_order)
          for (int a = 0; a < 3; ++a) p.positions[3 * isp + a] = p.positions[3 * (npos - 1) + a];
        p.positions.resize(3 * (npos - 1));
      }
      if (p.angle_next < p.angles.size() / 9) ++p.angle_next;
      --p.classes[(size_t)p.wait].remaining;
      p.wait = -1;
      ++p.inserted;
      *inserted = true;
    }
  }
  if (++p.counter == p.frequency) p.counter = 0;
  return GRAINS_OK;
}

int step_attempt(grains_gpu_system* s, Plan& p, bool* inserted) {
  if (p.preattempted) {
    p.preattempted = false;
    *inserted = false;
    return GRAINS_OK;
  }
  return attempt(s, p, inserted);
}

int insert_all(grains_gpu_system* s, Plan& p, long max_attempts) {
  long fails = 0;
  while (p.pending() > 0) {
    bool in = false;
    const bool tries = p.counter == 0;
    int e = attempt(s, p, &in);
    if (e) return e;
    if (in) {
      fails = 0;
      continue;
    }
    if (!tries) continue;
    if (++fails >= max_attempts) {
      char buf[160];
      std::snprintf(buf, sizeof(buf), "%ld particles could not be inserted (%ld attempts in a row failed)",
                    p.pending(), fails);
      s->last_error = buf;
      return GRAINS_CONFIG_ERROR;
    }
  }
  return GRAINS_OK;
}

}  // namespace grains_insert

extern "C" {

int grains_gpu_try_insert(grains_gpu_system* s, int* inserted) {
  if (!s || !inserted) return GRAINS_ARGUMENT_ERROR;
  *inserted = 0;
  if (!s->insertion || s->insertion->pending() == 0) return GRAINS_OK;
  if (!s->initialized) return GRAINS_ARGUMENT_ERROR;
  bool in = false;
  int e = grains_insert::step_attempt(s, *s->insertion, &in);
  *inserted = in ? 1 : 0;
  return e;
}

int grains_gpu_pending_insertions(grains_gpu_system* s, long* n) {
  if (!s || !n) return GRAINS_ARGUMENT_ERROR;
  *n = s->insertion ? s->insertion->pending() : 0;
  return GRAINS_OK;
}

}  // extern "C"
