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

// Options of <CollisionDetectionAlgorithm> (Grains.cpp:640-770;
// GrainsExec::m_colDetGJK_SV, m_colDetTolerance, m_colDetAcceleration,
// m_colDetBoundingVolume), set per system and installed here by each oracle
// entry point that evaluates contacts.
struct CollisionOptions {
  int sv = 0;         // 1: signed-volume sub-distance (GJK_SV.cpp)
  double tol = 1e-10; // relative tolerance of the optimality gap
  int accel = 0;      // 1: Nesterov acceleration
  int bv = 0;         // bounding-volume pre-test: 0 none, 1 OBB, 2 OBC
};
CollisionOptions g_col;

// Nesterov direction of closest_points (Convex.cpp:575-585, the same in
// GJK_SV.cpp:479-489)
inline V3 nesterovDirection(int it, V3 const& d, V3 const& v, V3 const& w) {
  double momentum = it / (it + 2.), omm = 1. - momentum;
  return momentum * d + momentum * omm * v + omm * omm * w;
}

// closest_points (Convex.cpp:536-627): Johnson's sub-distance, relative
// tolerance and acceleration of g_col
double gjkClosestPoints(ConvexRef const& a, ConvexRef const& b, Tf const& a2w, Tf const& b2w, V3& pa, V3& pb,
                        int& nbIter) {
  unsigned bits = 0, last = 0, last_bit = 0, all_bits = 0;
  Simplex S;
  int it = 0;
  const double relError = g_col.tol, absError = 1.e-4 * relError;
  bool acceleration = g_col.accel != 0;
  V3 v = applyTf(a2w, support(a, V3())) - applyTf(b2w, support(b, V3()));
  V3 w = v;
  V3 d = v;
  double dist = Norm(v);
  while (bits < 15 && dist > EPSILON2 && it < 1000) {
    ++it;
    last = 0;
    last_bit = 1;
    while (bits & last_bit) {
      ++last;
      last_bit <<= 1;
    }
    if (acceleration && it % 8 != 0) d = nesterovDirection(it, d, v, w);
    else d = v;
    S.p[last] = support(a, vecTimesBasis(-d, a2w));
    S.q[last] = support(b, vecTimesBasis(d, b2w));
    w = applyTf(a2w, S.p[last]) - applyTf(b2w, S.q[last]);
    double mu = dist - dot(v, w) / dist;
    if (mu < dist * relError || mu < absError) {
      if (!acceleration) break;
      if (Norm(d - v) < EPSILON) break;
      acceleration = false;
      S.p[last] = support(a, vecTimesBasis(-v, a2w));
      S.q[last] = support(b, vecTimesBasis(v, b2w));
      w = applyTf(a2w, S.p[last]) - applyTf(b2w, S.q[last]);
    }
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

// --- Signed-volume sub-distance (GJK_SV.cpp:1-420) --------------------------
// The subalgorithm of Montanari, Petrinic and Barbieri (2017) as the
// reference writes it: points selected by the bits of the simplex, signed
// volumes of the sub-simplices in the projection of largest area, and the
// recursion into the faces (edges) whose signed volume disagrees, keeping the
// closest.
inline bool svSameSign(double a, double b) { return !((a > 0) ^ (b > 0)); }

inline void svS1d(V3 const y[4], unsigned& bits, double (&lam)[4]) {
  bool s1_set = false;
  unsigned i1 = 0xffffffff, i2 = 0xffffffff;
  for (unsigned i = 0; i < 4; ++i)
    if (bits & (1u << i)) {
      if (s1_set) {
        i2 = i;
        break;
      }
      i1 = i;
      s1_set = true;
    }
  V3 t = y[i2] - y[i1];
  unsigned I = 0;
  double neg_tI = -t[0];
  if (fabs(t[1]) > fabs(neg_tI)) {
    I = 1;
    neg_tI = -t[1];
  }
  if (fabs(t[2]) > fabs(neg_tI)) {
    I = 2;
    neg_tI = -t[2];
  }
  double pI = dot(y[i2], t) / dot(t, t) * neg_tI + y[i2][I];
  double C[2] = {-y[i2][I] + pI, y[i1][I] - pI};
  bool sc0 = svSameSign(neg_tI, C[0]), sc1 = svSameSign(neg_tI, C[1]);
  if (sc0 && sc1) {
    lam[i1] = C[0] / neg_tI;
    lam[i2] = C[1] / neg_tI;
  } else if (sc0) {
    bits &= ~(1u << i2);
    lam[i1] = 1.;
  } else {
    bits &= ~(1u << i1);
    lam[i2] = 1.;
  }
}

inline void svS2d(V3 const y[4], unsigned& bits, double (&lam)[4]) {
  unsigned counter = 0, i0 = 0, i1 = 0, i2 = 0;
  for (unsigned i = 0; i < 4; ++i)
    if (bits & (1u << i)) {
      if (counter == 0) i0 = i;
      else if (counter == 1) i1 = i;
      else i2 = i;
      counter += 1;
    }
  V3 n = cross(y[i1] - y[i0], y[i2] - y[i0]);
  V3 p0 = (dot(y[i0], n) / dot(n, n)) * n;
  unsigned ix = 1, iy = 2;
  double mu_max = (y[i1][1] * y[i2][2] + y[i0][1] * y[i1][2] + y[i2][1] * y[i0][2] - y[i1][1] * y[i0][2] -
                   y[i2][1] * y[i1][2] - y[i0][1] * y[i2][2]);
  double mu = (y[i1][2] * y[i0][0] + y[i2][2] * y[i1][0] + y[i0][2] * y[i2][0] - y[i1][2] * y[i2][0] -
               y[i0][2] * y[i1][0] - y[i2][2] * y[i0][0]);
  if (fabs(mu) > fabs(mu_max)) {
    mu_max = mu;
    ix = 0;
  }
  mu = (y[i1][0] * y[i2][1] + y[i0][0] * y[i1][1] + y[i2][0] * y[i0][1] - y[i1][0] * y[i0][1] -
        y[i2][0] * y[i1][1] - y[i0][0] * y[i2][1]);
  if (fabs(mu) > fabs(mu_max)) {
    mu_max = mu;
    ix = 0;
    iy = 1;
  }
  double C[3];
  bool sc[3];
  C[0] = (p0[ix] * y[i1][iy] + p0[iy] * y[i2][ix] + y[i1][ix] * y[i2][iy] - p0[ix] * y[i2][iy] -
          p0[iy] * y[i1][ix] - y[i2][ix] * y[i1][iy]);
  sc[0] = svSameSign(mu_max, C[0]);
  C[1] = (p0[ix] * y[i2][iy] + p0[iy] * y[i0][ix] + y[i2][ix] * y[i0][iy] - p0[ix] * y[i0][iy] -
          p0[iy] * y[i2][ix] - y[i0][ix] * y[i2][iy]);
  sc[1] = svSameSign(mu_max, C[1]);
  C[2] = (p0[ix] * y[i0][iy] + p0[iy] * y[i1][ix] + y[i0][ix] * y[i1][iy] - p0[ix] * y[i1][iy] -
          p0[iy] * y[i0][ix] - y[i1][ix] * y[i0][iy]);
  sc[2] = svSameSign(mu_max, C[2]);
  if (sc[0] && sc[1] && sc[2]) {
    lam[i0] = C[0] / mu_max;
    lam[i1] = C[1] / mu_max;
    lam[i2] = C[2] / mu_max;
    return;
  }
  double d = 1.e9;
  unsigned new_bits = 0;
  const unsigned idx[3] = {i0, i1, i2};
  for (unsigned j = 0; j < 3; ++j) {
    if (sc[j]) continue;
    unsigned new_used = bits & ~(1u << idx[j]);
    double nl[4] = {0., 0., 0., 0.};
    svS1d(y, new_used, nl);
    V3 np(0., 0., 0.);
    for (unsigned i = 0; i < 4; ++i)
      if (new_used & (1u << i)) np += nl[i] * y[i];
    double d_star = dot(np, np);
    if (d_star < d) {
      new_bits = new_used;
      d = d_star;
      for (unsigned i = 0; i < 4; ++i) lam[i] = nl[i];
    }
  }
  bits = new_bits;
}

inline void svS3d(V3 const y[4], unsigned& bits, double (&lam)[4]) {
  double C[4];
  C[0] = y[3][0] * y[2][1] * y[1][2] + y[2][0] * y[1][1] * y[3][2] + y[1][0] * y[3][1] * y[2][2] -
         y[1][0] * y[2][1] * y[3][2] - y[2][0] * y[3][1] * y[1][2] - y[3][0] * y[1][1] * y[2][2];
  C[1] = y[0][0] * y[2][1] * y[3][2] + y[2][0] * y[3][1] * y[0][2] + y[3][0] * y[0][1] * y[2][2] -
         y[3][0] * y[2][1] * y[0][2] - y[2][0] * y[0][1] * y[3][2] - y[0][0] * y[3][1] * y[2][2];
  C[2] = y[3][0] * y[1][1] * y[0][2] + y[1][0] * y[0][1] * y[3][2] + y[0][0] * y[3][1] * y[1][2] -
         y[0][0] * y[1][1] * y[3][2] - y[1][0] * y[3][1] * y[0][2] - y[3][0] * y[0][1] * y[1][2];
  C[3] = y[0][0] * y[1][1] * y[2][2] + y[1][0] * y[2][1] * y[0][2] + y[2][0] * y[0][1] * y[1][2] -
         y[2][0] * y[1][1] * y[0][2] - y[1][0] * y[0][1] * y[2][2] - y[0][0] * y[2][1] * y[1][2];
  double dM = C[0] + C[1] + C[2] + C[3];
  bool sc[4];
  for (int i = 0; i < 4; ++i) sc[i] = svSameSign(dM, C[i]);
  if (sc[0] && sc[1] && sc[2] && sc[3]) {
    for (int i = 0; i < 4; ++i) lam[i] = C[i] / dM;
    return;
  }
  double d = 1.e9;
  unsigned new_bits = 0;
  for (unsigned j = 0; j < 4; ++j) {
    if (sc[j]) continue;
    unsigned new_used = bits & ~(1u << j);
    double nl[4] = {0., 0., 0., 0.};
    svS2d(y, new_used, nl);
    V3 np(0., 0., 0.);
    for (unsigned i = 0; i < 4; ++i)
      if (new_used & (1u << i)) np += nl[i] * y[i];
    double d_star = dot(np, np);
    if (d_star < d) {
      new_bits = new_used;
      d = d_star;
      for (unsigned i = 0; i < 4; ++i) lam[i] = nl[i];
    }
  }
  bits = new_bits;
}

// closest_points_GJK_SV (GJK_SV.cpp:428-525)
double gjkSVClosestPoints(ConvexRef const& a, ConvexRef const& b, Tf const& a2w, Tf const& b2w, V3& pa, V3& pb,
                          int& nbIter) {
  unsigned bits = 0, last = 0;
  V3 p[4], q[4], y[4];
  double lam[4] = {0., 0., 0., 0.};
  int it = 0;
  const double relError = g_col.tol, absError = 1.e-4 * relError;
  bool acceleration = g_col.accel != 0;
  V3 v = applyTf(a2w, support(a, V3())) - applyTf(b2w, support(b, V3()));
  V3 w = v, d = v;
  double dist = Norm(v);
  while (bits < 15 && dist > EPSILON2 && it < 1000) {
    ++it;
    for (unsigned ni = 0; ni < 4; ++ni)
      if (!(bits & (1u << ni))) {
        last = ni;
        break;
      }
    if (acceleration && it % 8 != 0) d = nesterovDirection(it, d, v, w);
    else d = v;
    p[last] = support(a, vecTimesBasis(-d, a2w));
    q[last] = support(b, vecTimesBasis(d, b2w));
    w = applyTf(a2w, p[last]) - applyTf(b2w, q[last]);
    double mu = dist - dot(v, w) / dist;
    if (mu < dist * relError || mu < absError) {
      if (!acceleration) break;
      if (Norm(d - v) < EPSILON) break;
      acceleration = false;
      p[last] = support(a, vecTimesBasis(-v, a2w));
      q[last] = support(b, vecTimesBasis(v, b2w));
      w = applyTf(a2w, p[last]) - applyTf(b2w, q[last]);
    }
    bool degen = false;
    for (unsigned i = 0; i < 4; ++i)
      if ((bits & (1u << i)) && y[i][0] == w[0] && y[i][1] == w[1] && y[i][2] == w[2]) degen = true;
    if (degen) break;
    y[last] = w;
    bits |= 1u << last;
    unsigned num = 0;
    for (unsigned i = 0; i < 4; ++i) num += (bits >> i) & 1u;
    if (num == 1) {
      for (unsigned i = 0; i < 4; ++i)
        if (bits & (1u << i)) lam[i] = 1.;
    } else if (num == 2) {
      svS1d(y, bits, lam);
    } else if (num == 3) {
      svS2d(y, bits, lam);
    } else {
      svS3d(y, bits, lam);
    }
    v = V3(0., 0., 0.);
    for (unsigned i = 0; i < 4; ++i)
      if (bits & (1u << i)) v += lam[i] * y[i];
    dist = Norm(v);
  }
  pa = V3(0., 0., 0.);
  pb = V3(0., 0., 0.);
  for (unsigned i = 0; i < 4; ++i)
    if (bits & (1u << i)) {
      pa += lam[i] * p[i];
      pb += lam[i] * q[i];
    }
  nbIter = it;
  return dist;
}

// --- Bounding-volume pre-tests (OBB.cpp:144-198, OBC.cpp:1-363) ------------
// OBB half extents of a convex at the identity (Box.cpp:1031-1056,
// Polyhedron.cpp:776-820, Sphere.cpp:528-537): the box's, the largest |x|,
// |y|, |z| of the polyhedron's vertices, the sphere's radius.
inline V3 obbExtent(ConvexRef const& c) {
  if (c.type == SH_SPHERE) return V3(c.r, c.r, c.r);
  if (c.type == SH_BOX) return c.e;
  V3 e(0., 0., 0.);
  for (V3 const& pt : c.sh->verts)
    for (int a = 0; a < 3; ++a) e[a] = fabs(pt[a]) > e[a] ? fabs(pt[a]) : e[a];
  return e;
}

// Separating-axis test of two OBBs (OBB.cpp:144-198): the 15 axes, with the
// 1e-6 (LOWEPS) margin on the absolute rotation
inline bool obbOverlap(V3 const& a, Tf const& a2w, V3 const& b, Tf const& b2w) {
  const double LOWEPS = 1.0e-6;
  double ori[3][3], oriAbs[3][3];
  V3 dc = b2w.o - a2w.o;
  // ori = transpose(Ba) * Bb ; cen = transpose(Ba) * (ob - oa)
  V3 cen;
  for (int i = 0; i < 3; ++i) cen[i] = a2w.B[0][i] * dc[0] + a2w.B[1][i] * dc[1] + a2w.B[2][i] * dc[2];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      ori[i][j] = a2w.B[0][i] * b2w.B[0][j] + a2w.B[1][i] * b2w.B[1][j] + a2w.B[2][i] * b2w.B[2][j];
      oriAbs[i][j] = fabs(ori[i][j]) + LOWEPS;
    }
  for (int i = 0; i < 3; ++i)
    if (fabs(cen[i]) > (a[i] + b[0] * oriAbs[i][0] + b[1] * oriAbs[i][1] + b[2] * oriAbs[i][2])) return false;
  for (int i = 0; i < 3; ++i)
    if (fabs(cen[0] * ori[0][i] + cen[1] * ori[1][i] + cen[2] * ori[2][i]) >
        (b[i] + a[0] * oriAbs[0][i] + a[1] * oriAbs[1][i] + a[2] * oriAbs[2][i]))
      return false;
  for (int j = 0; j < 3; ++j)
    for (int i = 0; i < 3; ++i)
      if (fabs(cen[(i + 2) % 3] * ori[(i + 1) % 3][j] - cen[(i + 1) % 3] * ori[(i + 2) % 3][j]) >
          (a[(i + 1) % 3] * oriAbs[(i + 2) % 3][j] + a[(i + 2) % 3] * oriAbs[(i + 1) % 3][j] +
           b[(j + 1) % 3] * oriAbs[i][(j + 2) % 3] + b[(j + 2) % 3] * oriAbs[i][(j + 1) % 3]))
        return false;
  return true;
}

// OBC of a convex (Box.cpp:1031-1056, Polyhedron.cpp:776-820,
// Sphere.cpp:528-537): axis along the OBB extent closest to another one,
// height twice that extent, radius the norm of the two others
struct OBCyl {
  double r, h;
  V3 e;
};
inline OBCyl obcOf(ConvexRef const& c) {
  if (c.type == SH_SPHERE) return OBCyl{c.r, 2. * c.r, V3(1., 0., 0.)};
  V3 ext = obbExtent(c);
  double xy = fabs(ext[0] - ext[1]), xz = fabs(ext[0] - ext[2]), yz = fabs(ext[1] - ext[2]);
  int zAxis = xy < xz ? 2 : 1;
  xy = xy < xz ? xy : xz;
  zAxis = xy < yz ? zAxis : 0;
  V3 e(0., 0., 0.);
  e[zAxis] = 1.;
  double h = 2. * ext[zAxis];
  double r = std::sqrt(ext[0] * ext[0] + ext[1] * ext[1] + ext[2] * ext[2] - ext[zAxis] * ext[zAxis]);
  return OBCyl{r, h, e};
}

inline int sgnOf(double v) { return (0. < v) - (v < 0.); }

// Real roots of x^4 + b x^3 + c x^2 + d x + e (OBC.cpp:143-222): depressed
// quartic, biquadratic case or Ferrari's resolvent cubic
inline void quarticRoots(double b, double c, double d, double e, double sol[4], int& nb) {
  nb = 0;
  const double b2 = b * b;
  const double p = c - 3. * b2 / 8.;
  const double q = b2 * b / 8. - b * c / 2. + d;
  const double r = -3. * b2 * b2 / 256. + e - b * d / 4. + b2 * c / 16.;
  const double p2 = p * p;
  if (fabs(q) < EPSILON) {
    const double del = p2 / 4. - r;
    if (del < 0.) return;
    const double m1 = -p / 2. + std::sqrt(del), m2 = -p / 2. - std::sqrt(del);
    if (m1 > 0.) {
      sol[nb++] = std::sqrt(m1) - b / 4.;
      sol[nb++] = -std::sqrt(m1) - b / 4.;
    }
    if (m2 > 0.) {
      sol[nb++] = std::sqrt(m2) - b / 4.;
      sol[nb++] = -std::sqrt(m2) - b / 4.;
    }
    return;
  }
  const double u = -p2 / 36. - r / 3.;
  const double v = -p2 * p / 216. + r * p / 6. - q * q / 16.;
  const double del = u * u * u + v * v;
  double m;
  if (del < 0) {
    m = 2. * std::sqrt(-u) * std::cos(std::acos(v / std::sqrt(-u) / u) / 3.) - p / 3.;
  } else {
    m = std::cbrt(-v + std::sqrt(del));
    m = m - u / m - p / 3.;
  }
  if (m < 0.) return;
  const double sm = std::sqrt(m / 2.);
  const double f1 = -p / 2. - m / 2. - q / sm / 4.;
  const double f2 = f1 + q / sm / 2.;
  if (f1 > 0.) {
    sol[nb++] = sm + std::sqrt(f1) - b / 4.;
    sol[nb++] = sm - std::sqrt(f1) - b / 4.;
  }
  if (f2 > 0.) {
    sol[nb++] = -sm + std::sqrt(f2) - b / 4.;
    sol[nb++] = -sm - std::sqrt(f2) - b / 4.;
  }
}

inline V3 basisTimes(Tf const& t, V3 const& v) {
  return V3(t.B[0][0] * v[0] + t.B[0][1] * v[1] + t.B[0][2] * v[2],
            t.B[1][0] * v[0] + t.B[1][1] * v[1] + t.B[1][2] * v[2],
            t.B[2][0] * v[0] + t.B[2][1] * v[1] + t.B[2][2] * v[2]);
}
inline V3 unitOf(V3 const& v) { return v / Norm(v); }

// One ellipse-circle step of the OBC test (OBC.cpp:300-355): the other
// cylinder's cap seen as an ellipse against this one's circle of radius rc
inline bool obcEllipseApart(double x0, double yTop, double yBottom, double ez, double ey, double re, double rc) {
  if (std::signbit(yTop) != std::signbit(yBottom)) return false;
  const double y0 = fabs(yTop) > fabs(yBottom) ? yBottom : yTop;
  if (!(x0 * x0 + y0 * y0 / ez / ez > re * re)) return false;
  const double ey_sq = ey * ey;
  const double A = re * re * ey_sq * ey_sq;
  const double B = 2. * y0 * ez / re / ey_sq;
  const double C = x0 * x0 / A;
  const double D = y0 * y0 * ez * ez / A;
  double sint[4];
  int nb;
  quarticRoots(-B, C + D - 1., B, -D, sint, nb);
  for (int i = 0; i < nb; ++i)
    if (std::signbit(sint[i]) != std::signbit(B)) {
      const double cost = -sgnOf(x0) * std::sqrt(1. - sint[i] * sint[i]);
      const double ptX = x0 + re * cost;
      const double ptY = y0 + re * ez * sint[i];
      if (ptX * ptX + ptY * ptY > rc * rc) return true;
    }
  return false;
}

// isContactBVolume(OBC, OBC) (OBC.cpp:226-363)
inline bool obcOverlap(OBCyl const& A, Tf const& a2w, OBCyl const& Bc, Tf const& b2w) {
  const double r1 = A.r, r2 = Bc.r, h1 = A.h / 2., h2 = Bc.h / 2.;
  const V3 eZ = basisTimes(a2w, A.e);
  const V3 e2 = basisTimes(b2w, Bc.e);
  const V3 xPt = b2w.o - a2w.o;
  const V3 eX = unitOf(cross(e2, eZ));
  const V3 eY = unitOf(cross(eZ, eX));
  const double x = dot(eX, xPt), y = dot(eY, xPt), z = dot(eZ, xPt);
  const double ey = dot(eY, e2), ez = dot(eZ, e2);
  if (fabs(x) >= r1 + r2) return false;
  {
    const double s2 = y / ey;
    const double s1 = z + s2 * ez;
    if (fabs(s1) < h1 && fabs(s2) < h2) return true;
  }
  {
    const double fey = fabs(ey) + 1.e-5, fez = fabs(ez) + 1.e-5;
    if (fabs(y) > r1 + h2 * fey + r2 * fez) return false;
    if (fabs(z) > h1 + h2 * fez + r2 * fey) return false;
    if (fabs(y * ez - z * ey) > r2 + h1 * fey + r1 * fez) return false;
    if (fabs(y * ey + z * ez) > h2 + h1 * fez + r1 * fey) return false;
  }
  if (fabs(ez) < 0.99 && fabs(ez) > 0.01) {
    if (obcEllipseApart(x, y + h2 * ey, y - h2 * ey, ez, ey, r2, r1)) return false;
    const V3 eY2 = unitOf(cross(eX, e2));
    const double y2 = -dot(eY2, xPt);
    const double ey2 = dot(eY2, eZ);
    if (obcEllipseApart(-x, y2 + h1 * ey2, y2 - h1 * ey2, ez, ey2, r1, r2)) return false;
  }
  return true;
}

// RigidBodyWithCrust::isContactBVolume (RigidBodyWithCrust.cpp:1095-1100)
// with the full transforms; bv = 1 OBB, 2 OBC
inline bool bvOverlap(int bv, ConvexRef const& a, Tf const& a2w, ConvexRef const& b, Tf const& b2w) {
  if (bv == 1) return obbOverlap(obbExtent(a), a2w, obbExtent(b), b2w);
  return obcOverlap(obcOf(a), a2w, obcOf(b), b2w);
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
