// Convex shape classes of non-spherical particles, built on the host: volume,
// inertia tensor per unit density, circumscribed radius, crust-shrink
// factors and the vertex table of polyhedra.
//
//   Sphere ........... Sphere.cpp:61-81, 152-155
//   Box .............. Box.cpp:181-230, 273-279
//   Polyhedron ....... Polyhedron.cpp:169-210, 324-359 with the tetrahedron
//                      contributions of GrainsExec::computeVolumeInertiaContrib
//                      (GrainsExec.cpp:746-797, F. Tonon, J. Math. Stat. 1 (2004))
//   crust scaling .... RigidBodyWithCrust.cpp:132-168 (bbox at the identity,
//                      Convex.cpp:29-41, BBox.cpp:235-239)
//   quaternion ....... Quaternion::setQuaternion(Matrix), Quaternion.cpp:200-255
#include <cmath>
#include <string>
#include <vector>

#include "dem_math.hpp"
#include "dem_shapes.hpp"

namespace dem {

namespace {

// Contribution of the tetrahedron (0, a, b, c) to the volume and to the
// inertia tensor (xx, xy, xz, yy, yz, zz) about the origin
void tetra(const double* a, const double* b, const double* c, double& vol, double* I) {
  const double x1 = 0., y1 = 0., z1 = 0.;
  const double x2 = a[0], y2 = a[1], z2 = a[2];
  const double x3 = b[0], y3 = b[1], z3 = b[2];
  const double x4 = c[0], y4 = c[1], z4 = c[2];
  const double det = std::fabs((x2 - x1) * (y3 - y1) * (z4 - z1) + (y2 - y1) * (z3 - z1) * (x4 - x1) +
                               (z2 - z1) * (x3 - x1) * (y4 - y1) - (z2 - z1) * (y3 - y1) * (x4 - x1) -
                               (x2 - x1) * (z3 - z1) * (y4 - y1) - (y2 - y1) * (x3 - x1) * (z4 - z1));
  vol += det / 6.;
  // Tonon's sums, each evaluated left to right as one expression
  I[0] += det *
          (y1 * y1 + y1 * y2 + y2 * y2 + y1 * y3 + y2 * y3 + y3 * y3 + y1 * y4 + y2 * y4 + y3 * y4 + y4 * y4 +
           z1 * z1 + z1 * z2 + z2 * z2 + z1 * z3 + z2 * z3 + z3 * z3 + z1 * z4 + z2 * z4 + z3 * z4 + z4 * z4) /
          60.;
  I[1] -= det *
          (2. * x1 * z1 + x2 * z1 + x3 * z1 + x4 * z1 + x1 * z2 + 2. * x2 * z2 + x3 * z2 + x4 * z2 + x1 * z3 +
           x2 * z3 + 2. * x3 * z3 + x4 * z3 + x1 * z4 + x2 * z4 + x3 * z4 + 2. * x4 * z4) /
          120.;
  I[2] -= det *
          (2. * x1 * y1 + x2 * y1 + x3 * y1 + x4 * y1 + x1 * y2 + 2. * x2 * y2 + x3 * y2 + x4 * y2 + x1 * y3 +
           x2 * y3 + 2. * x3 * y3 + x4 * y3 + x1 * y4 + x2 * y4 + x3 * y4 + 2. * x4 * y4) /
          120.;
  I[3] += det *
          (x1 * x1 + x1 * x2 + x2 * x2 + x1 * x3 + x2 * x3 + x3 * x3 + x1 * x4 + x2 * x4 + x3 * x4 + x4 * x4 +
           z1 * z1 + z1 * z2 + z2 * z2 + z1 * z3 + z2 * z3 + z3 * z3 + z1 * z4 + z2 * z4 + z3 * z4 + z4 * z4) /
          60.;
  I[4] -= det *
          (2. * y1 * z1 + y2 * z1 + y3 * z1 + y4 * z1 + y1 * z2 + 2. * y2 * z2 + y3 * z2 + y4 * z2 + y1 * z3 +
           y2 * z3 + 2. * y3 * z3 + y4 * z3 + y1 * z4 + y2 * z4 + y3 * z4 + 2. * y4 * z4) /
          120.;
  I[5] += det *
          (x1 * x1 + x1 * x2 + x2 * x2 + x1 * x3 + x2 * x3 + x3 * x3 + x1 * x4 + x2 * x4 + x3 * x4 + x4 * x4 +
           y1 * y1 + y1 * y2 + y2 * y2 + y1 * y3 + y2 * y3 + y3 * y3 + y1 * y4 + y2 * y4 + y3 * y4 + y4 * y4) /
          60.;
}

}  // namespace

int build_shape(int type, double radius, const double* L, int nv, const double* verts, int nf,
                const int* face_sizes, const int* face_idx, double crust, ShapeBuild& out, std::string& err) {
  ShapeDesc& d = out.desc;
  d.type = type;
  d.nv = 0;
  d.vert_off = 0;
  d.pad = 0;
  d.radius = 0.;
  d.crust = crust;
  for (int a = 0; a < 3; ++a) d.extent[a] = 0.;
  for (int i = 0; i < 6; ++i) out.inertia[i] = 0.;
  out.verts.clear();
  if (!(crust > 0.)) {
    err = "crust thickness must be positive";
    return 1;
  }
  double bext[3];
  if (type == kShapeSphere) {
    if (!(radius > 0.)) {
      err = "sphere radius must be positive";
      return 1;
    }
    d.radius = radius;
    out.volume = 4.0 * kPi * std::pow(radius, 3.) / 3.0;
    out.inertia[0] = out.inertia[3] = out.inertia[5] = 8.0 * kPi * std::pow(radius, 5.) / 15.0;
    d.circR = radius;
    bext[0] = bext[1] = bext[2] = radius;
    d.obb[0] = d.obb[1] = d.obb[2] = radius;  // Sphere::computeBVolume (Sphere.cpp:528-537)
  } else if (type == kShapeBox) {
    if (!L || !(L[0] > 0.) || !(L[1] > 0.) || !(L[2] > 0.)) {
      err = "box edge lengths must be positive";
      return 1;
    }
    const double ex = L[0] / 2., ey = L[1] / 2., ez = L[2] / 2.;
    d.extent[0] = ex;
    d.extent[1] = ey;
    d.extent[2] = ez;
    out.volume = 8.0 * ex * ey * ez;
    out.inertia[0] = 8.0 * ex * ey * ez * (ey * ey + ez * ez) / 3.0;
    out.inertia[3] = 8.0 * ex * ey * ez * (ex * ex + ez * ez) / 3.0;
    out.inertia[5] = 8.0 * ex * ey * ez * (ey * ey + ex * ex) / 3.0;
    d.circR = std::sqrt(ex * ex + ey * ey + ez * ez);
    bext[0] = ex;
    bext[1] = ey;
    bext[2] = ez;
    for (int a = 0; a < 3; ++a) d.obb[a] = d.extent[a];  // Box::computeBVolume (Box.cpp:1031-1056)
  } else if (type == kShapePoly) {
    if (nv < 4 || nf < 4 || !verts || !face_sizes || !face_idx) {
      err = "a polyhedron needs at least 4 vertices and 4 faces";
      return 1;
    }
    out.verts.assign(verts, verts + 3 * (size_t)nv);
    double vol = 0.;
    int off = 0;
    for (int f = 0; f < nf; ++f) {
      const int m = face_sizes[f];
      const int* fi = face_idx + off;
      off += m;
      if (m < 3) {
        err = "a polyhedron face needs at least 3 vertices";
        return 1;
      }
      for (int j = 0; j < m; ++j)
        if (fi[j] < 0 || fi[j] >= nv) {
          err = "polyhedron face index out of range";
          return 1;
        }
      if (m == 3) {
        tetra(verts + 3 * fi[0], verts + 3 * fi[1], verts + 3 * fi[2], vol, out.inertia);
      } else {
        // fan around the face centroid
        double G[3] = {0., 0., 0.};
        for (int j = 0; j < m; ++j)
          for (int a = 0; a < 3; ++a) G[a] += verts[3 * fi[j] + a];
        for (int a = 0; a < 3; ++a) G[a] /= double(m);
        for (int j = 0, k = m - 1; j < m; k = j++) tetra(G, verts + 3 * fi[k], verts + 3 * fi[j], vol, out.inertia);
      }
    }
    out.volume = vol;
    double ray = verts[0] * verts[0] + verts[1] * verts[1] + verts[2] * verts[2];
    for (int i = 1; i < nv; ++i) {
      const double* p = verts + 3 * i;
      double q = p[0] * p[0] + p[1] * p[1] + p[2] * p[2];
      if (q > ray) ray = q;
    }
    d.circR = std::sqrt(ray);
    d.nv = nv;
    // Polyhedron::computeBVolume (Polyhedron.cpp:776-820): the largest |x|,
    // |y|, |z| over the vertices, a box centred on the reference point
    for (int a = 0; a < 3; ++a) d.obb[a] = 0.;
    for (int i = 0; i < nv; ++i)
      for (int a = 0; a < 3; ++a) d.obb[a] = std::fabs(verts[3 * i + a]) > d.obb[a] ? std::fabs(verts[3 * i + a]) : d.obb[a];
    for (int a = 0; a < 3; ++a) {
      double mn = verts[a], mx = verts[a];
      for (int i = 1; i < nv; ++i) {
        mn = std::fmin(mn, verts[3 * i + a]);
        mx = std::fmax(mx, verts[3 * i + a]);
      }
      bext[a] = (mx - mn) / 2.;
    }
  } else {
    err = "unknown shape type";
    return 1;
  }
  for (int a = 0; a < 3; ++a) {
    d.bext[a] = bext[a];
    d.scaling[a] = bext[a] < kEps ? 1. : (bext[a] - crust) / bext[a];
  }
  return 0;
}

void quat_from_basis(const double* B, double* q) {
  const double r00 = B[0], r01 = B[1], r02 = B[2], r10 = B[3], r11 = B[4], r12 = B[5], r20 = B[6], r21 = B[7],
               r22 = B[8];
  double x = 0., y = 0., z = 0., w = 1., den;
  if (r11 > -r22 && r00 > -r11 && r00 > -r22) {
    den = std::pow(1. + r00 + r11 + r22, 0.5);
    w = 0.5 * den;
    x = 0.5 * (r21 - r12) / den;
    y = 0.5 * (r02 - r20) / den;
    z = 0.5 * (r10 - r01) / den;
  } else if (r11 < -r22 && r00 > r11 && r00 > r22) {
    den = std::pow(1. + r00 - r11 - r22, 0.5);
    w = 0.5 * (r21 - r12) / den;
    x = 0.5 * den;
    y = 0.5 * (r01 + r10) / den;
    z = 0.5 * (r20 + r02) / den;
  } else if (r11 > r22 && r00 < r11 && r00 < -r22) {
    den = std::pow(1. - r00 + r11 - r22, 0.5);
    w = 0.5 * (r02 - r20) / den;
    x = 0.5 * (r01 + r10) / den;
    y = 0.5 * den;
    z = 0.5 * (r12 + r21) / den;
  } else if (r11 < r22 && r00 < -r11 && r00 < r22) {
    den = std::pow(1. - r00 - r11 + r22, 0.5);
    w = 0.5 * (r10 - r01) / den;
    x = 0.5 * (r20 + r02) / den;
    y = 0.5 * (r12 + r21) / den;
    z = 0.5 * den;
  }
  q[0] = x;
  q[1] = y;
  q[2] = z;
  q[3] = w;
}

}  // namespace dem
