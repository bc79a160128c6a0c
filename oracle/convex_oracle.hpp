// TEST INFRASTRUCTURE (oracle).  Convex shapes other than spheres and the
// general narrow phase of the reference, restated in FP64 on the CPU:
// support functions, the Johnson-subalgorithm GJK distance, crust-shrunk
// transforms, the general ClosestPoint path with its ContactError handling,
// box / polyhedron volume and inertia, and the 3D angular acceleration.
// Included by grains_oracle.cpp inside its anonymous namespace (V3, Q, Part,
// Obs, Contact and EPSILON come from there).
//
// Parity notes: the Johnson subalgorithm, the termination tests and the
// witness-point formula follow Convex.cpp:219-627 step by step (same
// operation order), so the distance and witness points are pinned wherever
// the support points are unique.  Polyhedron support: see supportPoly.

enum ShapeType { SH_SPHERE = 0, SH_BOX = 1, SH_POLY = 2 };

struct Shape {
  int type;
  double radius;                          // sphere
  V3 extent;                              // box half lengths
  std::vector<V3> verts;                  // polyhedron vertices (local frame)
  std::vector<std::vector<int>> faces;    // polyhedron faces (vertex lists)
  std::vector<std::vector<int>> cobound;  // vertex adjacency (Polyhedron.cpp:324-352)
  int start_vertex;                       // warm start after the bbox computation
  double crust;
  double circR;        // circumscribed radius
  double volume;       // unit density
  double inertia[6];   // unit density: xx, xy, xz, yy, yz, zz
  V3 scaling;          // crust shrink factors (RigidBodyWithCrust.cpp:152-158)
  V3 bext;             // bounding-box half extents at the identity transform
};

// Sphere::support (Sphere.cpp:99-110)
inline V3 supportSphere(double r, V3 const& v) {
  double s = Norm(v);
  if (s > EPSILON) {
    double f = r / s;
    return V3(v[0] * f, v[1] * f, v[2] * f);
  }
  return V3();
}
// Box::support (Box.cpp:307-316)
inline V3 supportBox(V3 const& e, V3 const& v) {
  double n = Norm(v);
  if (n < EPSILON) return V3();
  return V3(v[0] < 0. ? -e[0] : e[0], v[1] < 0. ? -e[1] : e[1], v[2] < 0. ? -e[2] : e[2]);
}
// Polyhedron support as Polyhedron::support20 (Polyhedron.cpp:431-439): the
// first vertex of maximal projection.  The reference hill-climbs from a
// per-instance warm start instead (Polyhedron.cpp:265-293, moving only on
// improvements above 1e-16); both return a maximising vertex, and differ
// only among vertices tied to within 1e-16 -- which GJK meets whenever it
// converges onto a face normal, e.g. a sphere resting on a face.  The first
// maximum is what the device computes; the warm-start history of the
// reference is path dependent and not reproduced (parity unpinned there).
inline V3 supportPoly(Shape const& sh, int& curr, V3 const& v) {
  (void)curr;
  double n = Norm(v);
  if (!(n > EPSILON)) return V3();
  int c = 0;
  double h = dot(sh.verts[0], v), d;
  for (size_t i = 1; i < sh.verts.size(); ++i)
    if ((d = dot(sh.verts[i], v)) > h) {
      c = (int)i;
      h = d;
    }
  return sh.verts[c];
}

// A convex in a GJK call: its shape, radius for spheres and the warm start
// of polyhedra
struct ConvexRef {
  int type;
  double r;
  V3 e;
  Shape const* sh;
  int* curr;
};
inline V3 support(ConvexRef const& c, V3 const& v) {
  if (c.type == SH_SPHERE) return supportSphere(c.r, v);
  if (c.type == SH_BOX) return supportBox(c.e, v);
  return supportPoly(*c.sh, *c.curr, v);
}

// Transform = basis rows + origin
struct Tf {
  double B[3][3];
  V3 o;
};
// Transform::operator() (Transform.cpp:312-317)
inline V3 applyTf(Tf const& t, V3 const& p) {
  return V3(dot(V3(t.B[0][0], t.B[0][1], t.B[0][2]), p) + t.o[0],
            dot(V3(t.B[1][0], t.B[1][1], t.B[1][2]), p) + t.o[1],
            dot(V3(t.B[2][0], t.B[2][1], t.B[2][2]), p) + t.o[2]);
}
// Vector3 * Matrix (Matrix.cpp:238-241, 527-530)
inline V3 vecTimesBasis(V3 const& v, Tf const& t) {
  V3 r;
  for (int i = 0; i < 3; ++i) r[i] = t.B[0][i] * v[0] + t.B[1][i] * v[1] + t.B[2][i] * v[2];
  return r;
}

// --- Johnson sub-distance algorithm (Convex.cpp:219-435) --------------------
const double EPSILON2 = 1.e-15;
const double EPSILON3 = 1.e-20;

struct Simplex {
  V3 y[4], p[4], q[4];
  double det[16][4];
  double dp[4][4];
  Simplex() {
    for (int i = 0; i < 16; ++i)
      for (int j = 0; j < 4; ++j) det[i][j] = 0.;
    for (int i = 0; i < 4; ++i)
      for (int j = 0; j < 4; ++j) dp[i][j] = 0.;
  }
};

// sub-determinants of the subsets that contain `last`
inline void jDet(Simplex& S, unsigned bits, unsigned last, unsigned last_bit, unsigned all_bits) {
  for (unsigned i = 0, bit = 1; i < 4; ++i, bit <<= 1)
    if (bits & bit) S.dp[i][last] = S.dp[last][i] = dot(S.y[i], S.y[last]);
  S.dp[last][last] = dot(S.y[last], S.y[last]);
  S.det[last_bit][last] = 1.;
  for (unsigned j = 0, sj = 1; j < 4; ++j, sj <<= 1) {
    if (!(bits & sj)) continue;
    unsigned s2 = sj | last_bit;
    S.det[s2][j] = S.dp[last][last] - S.dp[last][j];
    S.det[s2][last] = S.dp[j][j] - S.dp[j][last];
    for (unsigned k = 0, sk = 1; k < j; ++k, sk <<= 1) {
      if (!(bits & sk)) continue;
      unsigned s3 = sk | s2;
      S.det[s3][k] = S.det[s2][j] * (S.dp[j][j] - S.dp[j][k]) + S.det[s2][last] * (S.dp[last][j] - S.dp[last][k]);
      S.det[s3][j] = S.det[sk | last_bit][k] * (S.dp[k][k] - S.dp[k][j]) +
                     S.det[sk | last_bit][last] * (S.dp[last][k] - S.dp[last][j]);
      S.det[s3][last] = S.det[sk | sj][k] * (S.dp[k][k] - S.dp[k][last]) +
                        S.det[sk | sj][j] * (S.dp[j][k] - S.dp[j][last]);
    }
  }
  if (all_bits == 15) {
    S.det[15][0] = S.det[14][1] * (S.dp[1][1] - S.dp[1][0]) + S.det[14][2] * (S.dp[2][1] - S.dp[2][0]) +
                   S.det[14][3] * (S.dp[3][1] - S.dp[3][0]);
    S.det[15][1] = S.det[13][0] * (S.dp[0][0] - S.dp[0][1]) + S.det[13][2] * (S.dp[2][0] - S.dp[2][1]) +
                   S.det[13][3] * (S.dp[3][0] - S.dp[3][1]);
    S.det[15][2] = S.det[11][0] * (S.dp[0][0] - S.dp[0][2]) + S.det[11][1] * (S.dp[1][0] - S.dp[1][2]) +
                   S.det[11][3] * (S.dp[3][0] - S.dp[3][2]);
    S.det[15][3] = S.det[7][0] * (S.dp[0][0] - S.dp[0][3]) + S.det[7][1] * (S.dp[1][0] - S.dp[1][3]) +
                   S.det[7][2] * (S.dp[2][0] - S.dp[2][3]);
  }
}
inline bool jValid(Simplex const& S, unsigned s, unsigned all_bits) {
  for (unsigned i = 0, bit = 1; i < 4; ++i, bit <<= 1) {
    if (!(all_bits & bit)) continue;
    if (s & bit) {
      if (S.det[s][i] <= EPSILON3) return false;
    } else if (S.det[s | bit][i] > 0.) {
      return false;
    }
  }
  return true;
}
inline void jVector(Simplex const& S, unsigned bits, V3& v) {
  double sum = 0.;
  v = V3(0., 0., 0.);
  for (unsigned i = 0, bit = 1; i < 4; ++i, bit <<= 1)
    if (bits & bit) {
      sum += S.det[bits][i];
      v += S.det[bits][i] * S.y[i];
    }
  v *= 1. / sum;
}
inline void jPoints(Simplex const& S, unsigned bits, V3& p1, V3& p2) {
  double sum = 0.;
  p1 = V3(0., 0., 0.);
  p2 = V3(0., 0., 0.);
  for (unsigned i = 0, bit = 1; i < 4; ++i, bit <<= 1)
    if (bits & bit) {
      sum += S.det[bits][i];
      p1 += S.det[bits][i] * S.p[i];
      p2 += S.det[bits][i] * S.q[i];
    }
  double s = 1. / sum;
  p1 *= s;
  p2 *= s;
}
inline bool jProper(Simplex const& S, unsigned s) {
  for (unsigned i = 0, bit = 1; i < 4; ++i, bit <<= 1)
    if ((s & bit) && S.det[s][i] <= EPSILON3) return false;
  return true;
}
// closest sub-simplex; false when the backup procedure was needed
inline bool jClosest(Simplex& S, unsigned& bits, unsigned last, unsigned last_bit, unsigned all_bits, V3& v) {
  jDet(S, bits, last, last_bit, all_bits);
  for (unsigned s = bits; s; --s) {
    if ((s & bits) == s && jValid(S, s | last_bit, all_bits)) {
      bits = s | last_bit;
      jVector(S, bits, v);
      return true;
    }
  }
  if (jValid(S, last_bit, all_bits)) {
    bits = last_bit;
    v = S.y[last];
    return true;
  }
  double min_dist2 = INFINITY;
  for (unsigned s = all_bits; s; --s) {
    if ((s & all_bits) == s && jProper(S, s)) {
      V3 u;
      jVector(S, s, u);
      double d2 = dot(u, u);
      if (d2 < min_dist2) {
        min_dist2 = d2;
        bits = s;
        v = u;
      }
    }
  }
  return false;
}
inline bool jDegenerate(Simplex const& S, unsigned all_bits, V3 const& w) {
  for (unsigned i = 0, bit = 1; i < 4; ++i, bit <<= 1)
    if ((all_bits & bit) && S.y[i][0] == w[0] && S.y[i][1] == w[1] && S.y[i][2] == w[2]) return true;
  return false;
}

// closest_points (Convex.cpp:536-627) without Nesterov acceleration (the
// default), relative tolerance 1e-10 (GrainsExec::m_colDetTolerance)
double gjkClosestPoints(ConvexRef const& a, ConvexRef const& b, Tf const& a2w, Tf const& b2w, V3& pa, V3& pb,
                        int& nbIter) {
  unsigned bits = 0, last = 0, last_bit = 0, all_bits = 0;
  Simplex S;
  int it = 0;
  const double relError = EPSILON, absError = 1.e-4 * relError;
  V3 v = applyTf(a2w, support(a, V3())) - applyTf(b2w, support(b, V3()));
  V3 w = v;
  double dist = Norm(v);
  while (bits < 15 && dist > EPSILON2 && it < 1000) {
    ++it;
    last = 0;
    last_bit = 1;
    while (bits & last_bit) {
      ++last;
      last_bit <<= 1;
    }
    V3 d = v;
    S.p[last] = support(a, vecTimesBasis(-d, a2w));
    S.q[last] = support(b, vecTimesBasis(d, b2w));
    w = applyTf(a2w, S.p[last]) - applyTf(b2w, S.q[last]);
    double mu = dist - dot(v, w) / dist;
    if (mu < dist * relError || mu < absError) break;
    if (jDegenerate(S, all_bits, w)) break;
    S.y[last] = w;
    all_bits = bits | last_bit;
    if (!jClosest(S, bits, last, last_bit, all_bits, v)) break;
    dist = Norm(v);
  }
  jPoints(S, bits, pa, pb);
  nbIter = it;
  return dist;
}

// --- Shapes: construction ---------------------------------------------------

// GrainsExec::computeVolumeInertiaContrib (GrainsExec.cpp:746-797), Tonon's
// tetrahedron formulas with the first vertex at the origin
void tetraContrib(V3 const& A2, V3 const& A3, V3 const& A4, double& vol, double* I) {
  double x1 = 0., x2 = A2[0], x3 = A3[0], x4 = A4[0];
  double y1 = 0., y2 = A2[1], y3 = A3[1], y4 = A4[1];
  double z1 = 0., z2 = A2[2], z3 = A3[2], z4 = A4[2];
  double det = std::fabs((x2 - x1) * (y3 - y1) * (z4 - z1) + (y2 - y1) * (z3 - z1) * (x4 - x1) +
                         (z2 - z1) * (x3 - x1) * (y4 - y1) - (z2 - z1) * (y3 - y1) * (x4 - x1) -
                         (x2 - x1) * (z3 - z1) * (y4 - y1) - (y2 - y1) * (x3 - x1) * (z4 - z1));
  vol += det / 6.;
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

// Shape bounding box at the identity transform (Convex.cpp:29-41,
// BBox.cpp:235-239) -> crust shrink factors (RigidBodyWithCrust.cpp:152-158)
void finishShape(Shape& sh) {
  int curr = sh.start_vertex;
  ConvexRef c{sh.type, sh.radius, sh.extent, &sh, &curr};
  V3 mn, mx;
  for (int a = 0; a < 3; ++a) {
    V3 e;
    e[a] = 1.;
    mn[a] = dot(e, support(c, -e));
  }
  for (int a = 0; a < 3; ++a) {
    V3 e;
    e[a] = 1.;
    mx[a] = dot(e, support(c, e));
  }
  V3 ext = (mx - mn) / 2.;
  for (int a = 0; a < 3; ++a) sh.scaling[a] = ext[a] < EPSILON ? 1. : (ext[a] - sh.crust) / ext[a];
  sh.bext = ext;
  sh.start_vertex = curr;
}

// Polyhedron::BuildPolyhedron (Polyhedron.cpp:324-359): adjacency from the
// face cycles, volume and inertia from the face fans
void buildPolyhedron(Shape& sh) {
  size_t nv = sh.verts.size();
  sh.cobound.assign(nv, std::vector<int>());
  sh.volume = 0.;
  for (int i = 0; i < 6; ++i) sh.inertia[i] = 0.;
  for (auto const& f : sh.faces) {
    size_t m = f.size();
    for (size_t j = 0, k = m - 1; j < m; k = j++) sh.cobound[f[k]].push_back(f[j]);
    if (m == 3) {
      tetraContrib(sh.verts[f[0]], sh.verts[f[1]], sh.verts[f[2]], sh.volume, sh.inertia);
    } else {
      V3 G;
      for (size_t j = 0; j < m; ++j) G += sh.verts[f[j]];
      G /= double(m);
      for (size_t j = 0, k = m - 1; j < m; k = j++)
        tetraContrib(G, sh.verts[f[k]], sh.verts[f[j]], sh.volume, sh.inertia);
    }
  }
  sh.start_vertex = 0;
  while (sh.cobound[sh.start_vertex].empty()) ++sh.start_vertex;
  // Polyhedron::computeCircumscribedRadius (Polyhedron.cpp:201-210)
  double ray = dot(sh.verts[0], sh.verts[0]), d;
  for (size_t i = 1; i < nv; ++i)
    if ((d = dot(sh.verts[i], sh.verts[i])) > ray) ray = d;
  sh.circR = std::sqrt(ray);
}

// Box volume / inertia / circumscribed radius (Box.cpp:181-230, 273-279)
void buildBoxShape(Shape& sh) {
  V3 const& e = sh.extent;
  sh.volume = 8.0 * e[0] * e[1] * e[2];
  sh.inertia[1] = sh.inertia[2] = sh.inertia[4] = 0.;
  sh.inertia[0] = 8.0 * e[0] * e[1] * e[2] * (e[1] * e[1] + e[2] * e[2]) / 3.0;
  sh.inertia[3] = 8.0 * e[0] * e[1] * e[2] * (e[0] * e[0] + e[2] * e[2]) / 3.0;
  sh.inertia[5] = 8.0 * e[0] * e[1] * e[2] * (e[1] * e[1] + e[0] * e[0]) / 3.0;
  sh.circR = Norm(e);
  sh.start_vertex = 0;
}

// Transform::composeLeftByRotation(Quaternion) (Transform.cpp:140-159)
void composeLeftByRotation(double B[3][3], Q const& q) {
  double qx = q.v[0], qy = q.v[1], qz = q.v[2], qw = q.w;
  for (int i = 0; i < 3; ++i) {
    double px = qy * B[2][i] - qz * B[1][i] + qw * B[0][i];
    double py = qz * B[0][i] - qx * B[2][i] + qw * B[1][i];
    double pz = qx * B[1][i] - qy * B[0][i] + qw * B[2][i];
    double pw = -qx * B[0][i] - qy * B[1][i] - qz * B[2][i];
    B[0][i] = qy * pz - qz * py - pw * qx + qw * px;
    B[1][i] = qz * px - qx * pz - pw * qy + qw * py;
    B[2][i] = qx * py - qy * px - pw * qz + qw * pz;
  }
}
