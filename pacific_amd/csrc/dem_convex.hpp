// Non-spherical particles on the device: support functions, the GJK distance
// with Johnson's sub-distance algorithm, the general ClosestPoint path with
// crust-shrunk transforms and its ContactError handling, and the rigid-body
// kinematics of ParticleKinematics3D.
//
//   RigidBodyWithCrust::ClosestPoint ........ RigidBodyWithCrust.cpp:205-354
//   ClosestPoint_ErreurHandling ............. RigidBodyWithCrust.cpp:534-590, 770-798
//   closest_points (Johnson, no Nesterov) ... Convex.cpp:219-435, 536-627
//   Sphere / Box / Polyhedron::support ...... Sphere.cpp:99-110, Box.cpp:307-316,
//                                              Polyhedron.cpp:265-293, 431-439
//   computeAngularAccelerationBodyFixed ..... ParticleKinematics3D.cpp:56-77
//   Transform::composeLeftByRotation ........ Transform.cpp:140-159
//
// The polyhedron support scans all vertices and keeps the first maximum
// (Polyhedron::support20); the reference hill-climbs from a per-particle warm
// start.  Both return the maximising vertex; they can differ only among tied
// maxima, i.e. for a face exactly parallel to the search direction.
#pragma once
#include <type_traits>

#include "dem_math.hpp"
#include "dem_types.hpp"

namespace dem {

constexpr double kEps2 = 1.e-15;  // Basic.hh EPSILON2
constexpr double kEps3 = 1.e-20;  // Basic.hh EPSILON3

// A rigid body with crust as ClosestPoint sees it
struct ConvexGeo {
  int type, nv;
  const double* verts;
  double r;       // sphere radius
  d3 e;           // box half extents
  double B[9];    // basis rows of the Transform (no crust)
  d3 o;           // origin
  d3 scaling;     // crust shrink factors
  d3 bext;        // bbox half extents at the identity transform
  double crust, circR;
  d3 obb;         // OBB half extents at the identity (bounding-volume pre-test)
  GjkOpts opt;    // the system's narrow-phase options
};

struct Tf {
  double B[9];
  d3 o;
};

// Transform::operator() (Transform.cpp:312-317)
DEM_HD d3 tf_apply(const Tf& t, d3 p) {
  return d3{dot(d3{t.B[0], t.B[1], t.B[2]}, p) + t.o.x, dot(d3{t.B[3], t.B[4], t.B[5]}, p) + t.o.y,
            dot(d3{t.B[6], t.B[7], t.B[8]}, p) + t.o.z};
}
// Vector3 * Matrix (Matrix.cpp:238-241, 527-530)
DEM_HD d3 tf_vec_basis(d3 v, const Tf& t) {
  return d3{t.B[0] * v.x + t.B[3] * v.y + t.B[6] * v.z, t.B[1] * v.x + t.B[4] * v.y + t.B[7] * v.z,
            t.B[2] * v.x + t.B[5] * v.y + t.B[8] * v.z};
}
// Transform::composeWithScaling (Matrix.cpp:347-358)
DEM_HD Tf tf_scaled(const ConvexGeo& g, d3 sc) {
  Tf t;
  for (int i = 0; i < 3; ++i) {
    t.B[3 * i] = g.B[3 * i] * sc.x;
    t.B[3 * i + 1] = g.B[3 * i + 1] * sc.y;
    t.B[3 * i + 2] = g.B[3 * i + 2] * sc.z;
  }
  t.o = g.o;
  return t;
}

DEM_HD d3 support(const ConvexGeo& g, d3 v) {
  if (g.type == kShapeSphere) {
    double s = norm(v);
    if (s > kEps) {
      double f = g.r / s;
      return d3{v.x * f, v.y * f, v.z * f};
    }
    return d3{0., 0., 0.};
  }
  if (g.type == kShapeBox) {
    double n = norm(v);
    if (n < kEps) return d3{0., 0., 0.};
    return d3{v.x < 0. ? -g.e.x : g.e.x, v.y < 0. ? -g.e.y : g.e.y, v.z < 0. ? -g.e.z : g.e.z};
  }
  double n = norm(v);
  if (!(n > kEps)) return d3{0., 0., 0.};
  const double* p = g.verts;
  d3 best{p[0], p[1], p[2]};
  double h = dot(best, v);
  for (int i = 1; i < g.nv; ++i) {
    d3 c{p[3 * i], p[3 * i + 1], p[3 * i + 2]};
    double d = dot(c, v);
    if (d > h) {
      h = d;
      best = c;
    }
  }
  return best;
}

// --- Johnson sub-distance algorithm -----------------------------------------
// The simplex of up to four points of the Minkowski difference and the
// determinant cache of Convex.cpp:219-435, with the same arithmetic in the
// same order.  Every table index below is a compile-time constant -- the
// slot of the point being added is a template argument, the subsets are
// enumerated by unrolled loops -- so that the tables live in registers; the
// run-time masks only pick which of the constant-indexed branches runs.
// Only the 32 entries det[s][i] with i in s and the upper triangle of the
// symmetric dot-product table are ever read.
struct Simplex {
  d3 y[4], p[4], q[4];
  double det[16][4];
  double dp[10];  // y_i . y_j for i <= j
};

// index of y_i . y_j in Simplex::dp, a constant expression (the tables are
// addressed at compile time only)
template <int I, int J>
constexpr int kDp = I <= J ? I * 4 - I * (I - 1) / 2 + (J - I) : J * 4 - J * (J - 1) / 2 + (I - J);

// f(integral_constant<int, i>) for i = B, B+1, ..., E-1
template <int B, int E, class F>
DEM_HD void sfor(F&& f) {
  if constexpr (B < E) {
    f(std::integral_constant<int, B>{});
    sfor<B + 1, E>(f);
  }
}
// f(integral_constant<int, i>) for i = B, B-1, ..., E+1
template <int B, int E, class F>
DEM_HD void sfor_down(F&& f) {
  if constexpr (B > E) {
    f(std::integral_constant<int, B>{});
    sfor_down<B - 1, E>(f);
  }
}

// computeDet (Convex.cpp:219-268) for the point just added in slot LAST
template <int LAST>
DEM_HD void j_det(Simplex& S, unsigned bits, unsigned all_bits) {
  constexpr unsigned last_bit = 1u << LAST;
  sfor<0, 4>([&](auto I) {
    constexpr int i = decltype(I)::value;
    if constexpr (i != LAST)
      if (bits & (1u << i)) S.dp[kDp<i, LAST>] = dot(S.y[i], S.y[LAST]);
  });
  S.dp[kDp<LAST, LAST>] = dot(S.y[LAST], S.y[LAST]);
  S.det[last_bit][LAST] = 1.;
  sfor<0, 4>([&](auto J) {
    constexpr int j = decltype(J)::value;
    if constexpr (j != LAST) {
      constexpr unsigned sj = 1u << j, s2 = sj | last_bit;
      if (bits & sj) {
        S.det[s2][j] = S.dp[kDp<LAST, LAST>] - S.dp[kDp<LAST, j>];
        S.det[s2][LAST] = S.dp[kDp<j, j>] - S.dp[kDp<j, LAST>];
        sfor<0, j>([&](auto K) {
          constexpr int k = decltype(K)::value;
          if constexpr (k != LAST) {
            constexpr unsigned sk = 1u << k, s3 = sk | s2;
            if (bits & sk) {
              S.det[s3][k] = S.det[s2][j] * (S.dp[kDp<j, j>] - S.dp[kDp<j, k>]) +
                             S.det[s2][LAST] * (S.dp[kDp<LAST, j>] - S.dp[kDp<LAST, k>]);
              S.det[s3][j] = S.det[sk | last_bit][k] * (S.dp[kDp<k, k>] - S.dp[kDp<k, j>]) +
                             S.det[sk | last_bit][LAST] * (S.dp[kDp<LAST, k>] - S.dp[kDp<LAST, j>]);
              S.det[s3][LAST] = S.det[sk | sj][k] * (S.dp[kDp<k, k>] - S.dp[kDp<k, LAST>]) +
                                S.det[sk | sj][j] * (S.dp[kDp<j, k>] - S.dp[kDp<j, LAST>]);
            }
          }
        });
      }
    }
  });
  if (all_bits == 15) {
    const double* dp = S.dp;
    S.det[15][0] = S.det[14][1] * (dp[kDp<1, 1>] - dp[kDp<1, 0>]) + S.det[14][2] * (dp[kDp<2, 1>] - dp[kDp<2, 0>]) +
                   S.det[14][3] * (dp[kDp<3, 1>] - dp[kDp<3, 0>]);
    S.det[15][1] = S.det[13][0] * (dp[kDp<0, 0>] - dp[kDp<0, 1>]) + S.det[13][2] * (dp[kDp<2, 0>] - dp[kDp<2, 1>]) +
                   S.det[13][3] * (dp[kDp<3, 0>] - dp[kDp<3, 1>]);
    S.det[15][2] = S.det[11][0] * (dp[kDp<0, 0>] - dp[kDp<0, 2>]) + S.det[11][1] * (dp[kDp<1, 0>] - dp[kDp<1, 2>]) +
                   S.det[11][3] * (dp[kDp<3, 0>] - dp[kDp<3, 2>]);
    S.det[15][3] = S.det[7][0] * (dp[kDp<0, 0>] - dp[kDp<0, 3>]) + S.det[7][1] * (dp[kDp<1, 0>] - dp[kDp<1, 3>]) +
                   S.det[7][2] * (dp[kDp<2, 0>] - dp[kDp<2, 3>]);
  }
}
// valid (Convex.cpp:272-290) for the subset SS
template <unsigned SS>
DEM_HD bool j_valid(const Simplex& S, unsigned all_bits) {
  bool ok = true;
  sfor<0, 4>([&](auto I) {
    constexpr int i = decltype(I)::value;
    constexpr unsigned bit = 1u << i;
    if (!ok || !(all_bits & bit)) return;
    if constexpr ((SS & bit) != 0u) {
      if (S.det[SS][i] <= kEps3) ok = false;
    } else {
      if (S.det[SS | bit][i] > 0.) ok = false;
    }
  });
  return ok;
}
// computeVector (Convex.cpp:294-306)
template <unsigned SS>
DEM_HD d3 j_vector(const Simplex& S) {
  double sum = 0.;
  d3 v{0., 0., 0.};
  sfor<0, 4>([&](auto I) {
    constexpr int i = decltype(I)::value;
    if constexpr ((SS & (1u << i)) != 0u) {
      sum += S.det[SS][i];
      v += S.det[SS][i] * S.y[i];
    }
  });
  v *= 1. / sum;
  return v;
}
// computePoints (Convex.cpp:310-327)
template <unsigned SS>
DEM_HD void j_points_t(const Simplex& S, d3& p1, d3& p2) {
  double sum = 0.;
  p1 = d3{0., 0., 0.};
  p2 = d3{0., 0., 0.};
  sfor<0, 4>([&](auto I) {
    constexpr int i = decltype(I)::value;
    if constexpr ((SS & (1u << i)) != 0u) {
      sum += S.det[SS][i];
      p1 += S.det[SS][i] * S.p[i];
      p2 += S.det[SS][i] * S.q[i];
    }
  });
  double s = 1. / sum;
  p1 *= s;
  p2 *= s;
}
DEM_HD void j_points(const Simplex& S, unsigned bits, d3& p1, d3& p2) {
  bool done = false;
  sfor_down<15, -1>([&](auto B) {
    constexpr unsigned b = (unsigned)decltype(B)::value;
    if (!done && bits == b) {
      j_points_t<b>(S, p1, p2);
      done = true;
    }
  });
}
// proper (Convex.cpp:331-338)
template <unsigned SS>
DEM_HD bool j_proper(const Simplex& S) {
  bool ok = true;
  sfor<0, 4>([&](auto I) {
    constexpr int i = decltype(I)::value;
    if constexpr ((SS & (1u << i)) != 0u)
      if (S.det[SS][i] <= kEps3) ok = false;
  });
  return ok;
}
// closest (Convex.cpp:342-414): the smallest valid subset containing LAST,
// else the backup procedure over every proper subset
template <int LAST>
DEM_HD bool j_closest(Simplex& S, unsigned& bits, unsigned all_bits, d3& v) {
  constexpr unsigned last_bit = 1u << LAST;
  j_det<LAST>(S, bits, all_bits);
  bool found = false;
  sfor_down<15, 0>([&](auto SI) {
    constexpr unsigned s = (unsigned)decltype(SI)::value;
    if constexpr ((s & last_bit) == 0u) {
      if (!found && (s & bits) == s && j_valid<s | last_bit>(S, all_bits)) {
        bits = s | last_bit;
        v = j_vector<s | last_bit>(S);
        found = true;
      }
    }
  });
  if (found) return true;
  if (j_valid<last_bit>(S, all_bits)) {
    bits = last_bit;
    v = S.y[LAST];
    return true;
  }
  double min_dist2 = INFINITY;
  sfor_down<15, 0>([&](auto SI) {
    constexpr unsigned s = (unsigned)decltype(SI)::value;
    if ((s & all_bits) == s && j_proper<s>(S)) {
      d3 u = j_vector<s>(S);
      double d2 = dot(u, u);
      if (d2 < min_dist2) {
        min_dist2 = d2;
        bits = s;
        v = u;
      }
    }
  });
  return false;
}
// degenerate (Convex.cpp:418-435)
DEM_HD bool j_degenerate(const Simplex& S, unsigned all_bits, d3 w) {
  bool deg = false;
  sfor<0, 4>([&](auto I) {
    constexpr int i = decltype(I)::value;
    if ((all_bits & (1u << i)) && S.y[i].x == w.x && S.y[i].y == w.y && S.y[i].z == w.z) deg = true;
  });
  return deg;
}

// One iteration of closest_points with the new point in slot LAST: 1 to go
// on, 0 when the loop ends.
// sep >= 0: 2 as soon as the support point w proves the distance larger
// than sep (the lower bound v.w / |v| of the distance exceeds it, with a
// margin above the rounding of v.w), so that a pair that cannot be in
// contact stops at once; the distance itself is then not computed.
// Nesterov direction of closest_points (Convex.cpp:575-585; GJK_SV.cpp:479-489)
DEM_HD d3 nesterov_dir(int it, d3 d, d3 v, d3 w) {
  const double momentum = it / (it + 2.), omm = 1. - momentum;
  return momentum * d + momentum * omm * v + omm * omm * w;
}

// The search direction, support points and optimality-gap test of one
// iteration, with the tolerance and acceleration of o: 1 to go on with the
// support points P, Q (local frames) and w, 0 when the loop ends, 2 when w
// proves the distance larger than sep (sep >= 0, acceleration off).
DEM_HD int gjk_search(const ConvexGeo& A, const Tf& a2w, const ConvexGeo& B, const Tf& b2w, const GjkOpts& o,
                      int it, d3 v, double dist, double sep, d3& d, d3& w, bool& accel, d3& P, d3& Q) {
  const double relError = o.tol, absError = 1.e-4 * relError;
  d = (accel && it % 8 != 0) ? nesterov_dir(it, d, v, w) : v;
  P = support(A, tf_vec_basis(-d, a2w));
  Q = support(B, tf_vec_basis(d, b2w));
  w = tf_apply(a2w, P) - tf_apply(b2w, Q);
  if (sep >= 0. && !o.accel && dot(v, w) > sep * dist + 1.e-12 * dist * (fabs(w.x) + fabs(w.y) + fabs(w.z)))
    return 2;
  const double mu = dist - dot(v, w) / dist;
  if (mu < dist * relError || mu < absError) {
    if (!accel) return 0;
    if (norm(d - v) < kEps) return 0;
    accel = false;
    P = support(A, tf_vec_basis(-v, a2w));
    Q = support(B, tf_vec_basis(v, b2w));
    w = tf_apply(a2w, P) - tf_apply(b2w, Q);
  }
  return 1;
}

template <int LAST>
DEM_HD int gjk_iter(Simplex& S, const ConvexGeo& A, const Tf& a2w, const ConvexGeo& B, const Tf& b2w,
                    unsigned& bits, unsigned& all_bits, d3& v, double& dist, double sep, int it, d3& d, d3& w,
                    bool& accel) {
  constexpr unsigned last_bit = 1u << LAST;
  const int go = gjk_search(A, a2w, B, b2w, A.opt, it, v, dist, sep, d, w, accel, S.p[LAST], S.q[LAST]);
  if (go != 1) return go;
  if (j_degenerate(S, all_bits, w)) return 0;
  S.y[LAST] = w;
  all_bits = bits | last_bit;
  if (!j_closest<LAST>(S, bits, all_bits, v)) return 0;
  dist = norm(v);
  return 1;
}

// --- Signed-volume sub-distance (GJK_SV.cpp:1-420) --------------------------
// Montanari, Petrinic and Barbieri (2017): the closest point of the simplex
// from signed volumes of its sub-simplices in the projection of largest
// area, recursing into the faces (edges) whose volume disagrees in sign.
// The points are named by their slot bits as in Johnson's; every slot index
// is a template argument so that the simplex lives in registers.
// Barycentric weights of the four slots, as named scalars (an array of
// them, written through references, would sit in scratch memory).
struct L4 {
  double w0, w1, w2, w3;
};
template <int I>
DEM_HD double& lw(L4& l) {
  if constexpr (I == 0) return l.w0;
  else if constexpr (I == 1) return l.w1;
  else if constexpr (I == 2) return l.w2;
  else return l.w3;
}

struct SvSimplex {
  d3 y[4], p[4], q[4];
  L4 lam;
};

DEM_HD bool sv_same_sign(double a, double b) { return !((a > 0) ^ (b > 0)); }
// sum of lam_i y_i over the slots of bits, in slot order
DEM_HD d3 sv_point(const d3 (&y)[4], unsigned bits, const L4& l) {
  d3 v{0., 0., 0.};
  if (bits & 1u) v += l.w0 * y[0];
  if (bits & 2u) v += l.w1 * y[1];
  if (bits & 4u) v += l.w2 * y[2];
  if (bits & 8u) v += l.w3 * y[3];
  return v;
}

// component i of v (0, 1, 2) selected among the values
DEM_HD double sel3(int i, d3 v) { return i == 0 ? v.x : (i == 1 ? v.y : v.z); }
// c ? a : b component by component (a select of values: a select of the
// addresses of a and b would keep them in scratch memory)
DEM_HD d3 pick(bool c, d3 a, d3 b) { return d3{c ? a.x : b.x, c ? a.y : b.y, c ? a.z : b.z}; }
// a and b equal bit for bit (signed zeros told apart)
DEM_HD bool same_bits(d3 a, d3 b) {
  return __builtin_bit_cast(unsigned long long, a.x) == __builtin_bit_cast(unsigned long long, b.x) &&
         __builtin_bit_cast(unsigned long long, a.y) == __builtin_bit_cast(unsigned long long, b.y) &&
         __builtin_bit_cast(unsigned long long, a.z) == __builtin_bit_cast(unsigned long long, b.z);
}

// The sub-distance works on the active points gathered in slot order into
// named values (a, b, c, ...): one copy of the edge, triangle and
// tetrahedron code each, the faces (edges) to recurse into picked by a loop
// that is not unrolled.  The arithmetic is the slot-indexed one's, in the
// same order.  Each returns the mask of the points it keeps (bit k = k-th
// point) and writes their weights; a dropped point's weight is undefined.

// the point of the edge a-b closest to the origin (GJK_SV.cpp:19-93)
DEM_HD unsigned sv_edge(const d3& a, const d3& b, double& la, double& lb) {
  const d3 t = b - a;
  int I = 0;
  double neg_tI = -t.x;
  if (fabs(t.y) > fabs(neg_tI)) {
    I = 1;
    neg_tI = -t.y;
  }
  if (fabs(t.z) > fabs(neg_tI)) {
    I = 2;
    neg_tI = -t.z;
  }
  const double y2I = sel3(I, b), y1I = sel3(I, a);
  const double pI = dot(b, t) / dot(t, t) * neg_tI + y2I;
  const double C0 = -y2I + pI, C1 = y1I - pI;
  const bool s0 = sv_same_sign(neg_tI, C0), s1 = sv_same_sign(neg_tI, C1);
  const bool both = s0 && s1;
  la = both ? C0 / neg_tI : 1.;
  lb = both ? C1 / neg_tI : 1.;
  return both ? 3u : (s0 ? 1u : 2u);
}

// the triangle a-b-c (GJK_SV.cpp:99-228); the edges without a point whose
// signed area disagrees in sign are tried in point order, the closest kept
// (GJK_SV.cpp:190-225)
DEM_HD unsigned sv_tri(const d3& a, const d3& b, const d3& c, double& la, double& lb, double& lc) {
  const d3 n = cross(b - a, c - a);
  const d3 p0 = (dot(a, n) / dot(n, n)) * n;
  int ix = 1, iy = 2;
  double mu_max = (b.y * c.z + a.y * b.z + c.y * a.z - b.y * a.z - c.y * b.z - a.y * c.z);
  double mu = (b.z * a.x + c.z * b.x + a.z * c.x - b.z * c.x - a.z * b.x - c.z * a.x);
  if (fabs(mu) > fabs(mu_max)) {
    mu_max = mu;
    ix = 0;
  }
  mu = (b.x * c.y + a.x * b.y + c.x * a.y - b.x * a.y - c.x * b.y - a.x * c.y);
  if (fabs(mu) > fabs(mu_max)) {
    mu_max = mu;
    ix = 0;
    iy = 1;
  }
  const double px = sel3(ix, p0), py = sel3(iy, p0);
  const double y0x = sel3(ix, a), y0y = sel3(iy, a), y1x = sel3(ix, b), y1y = sel3(iy, b), y2x = sel3(ix, c),
               y2y = sel3(iy, c);
  const double C0 = (px * y1y + py * y2x + y1x * y2y - px * y2y - py * y1x - y2x * y1y);
  const double C1 = (px * y2y + py * y0x + y2x * y0y - px * y0y - py * y2x - y0x * y2y);
  const double C2 = (px * y0y + py * y1x + y0x * y1y - px * y1y - py * y0x - y1x * y0y);
  const bool s0 = sv_same_sign(mu_max, C0), s1 = sv_same_sign(mu_max, C1), s2 = sv_same_sign(mu_max, C2);
  if (s0 && s1 && s2) {
    la = C0 / mu_max;
    lb = C1 / mu_max;
    lc = C2 / mu_max;
    return 7u;
  }
  la = lb = lc = 0.;
  double dbest = 1.e9;
  unsigned best = 0;
#pragma nounroll
  for (int j = 0; j < 3; ++j) {
    if (j == 0 ? s0 : (j == 1 ? s1 : s2)) continue;
    // the edge without point j, its points in order
    const d3 e0 = pick(j == 0, b, a), e1 = pick(j == 2, b, c);
    double w0, w1;
    const unsigned k = sv_edge(e0, e1, w0, w1);
    d3 np{0., 0., 0.};
    if (k & 1u) np += w0 * e0;
    if (k & 2u) np += w1 * e1;
    const double ds = dot(np, np);
    if (ds < dbest) {
      dbest = ds;
      const unsigned m0 = j == 0 ? 2u : 1u, m1 = j == 2 ? 2u : 4u;
      best = ((k & 1u) ? m0 : 0u) | ((k & 2u) ? m1 : 0u);
      la = j == 0 ? la : w0;
      lb = j == 0 ? w0 : (j == 2 ? w1 : lb);
      lc = j == 2 ? lc : w1;
    }
  }
  return best;
}

// the tetrahedron y0-y1-y2-y3 (GJK_SV.cpp:234-373), its faces tried in
// point order as the triangle's edges (GJK_SV.cpp:338-370)
DEM_HD unsigned sv_tet(const d3 (&y)[4], L4& lam) {
  const double C0 = y[3].x * y[2].y * y[1].z + y[2].x * y[1].y * y[3].z + y[1].x * y[3].y * y[2].z -
                    y[1].x * y[2].y * y[3].z - y[2].x * y[3].y * y[1].z - y[3].x * y[1].y * y[2].z;
  const double C1 = y[0].x * y[2].y * y[3].z + y[2].x * y[3].y * y[0].z + y[3].x * y[0].y * y[2].z -
                    y[3].x * y[2].y * y[0].z - y[2].x * y[0].y * y[3].z - y[0].x * y[3].y * y[2].z;
  const double C2 = y[3].x * y[1].y * y[0].z + y[1].x * y[0].y * y[3].z + y[0].x * y[3].y * y[1].z -
                    y[0].x * y[1].y * y[3].z - y[1].x * y[3].y * y[0].z - y[3].x * y[0].y * y[1].z;
  const double C3 = y[0].x * y[1].y * y[2].z + y[1].x * y[2].y * y[0].z + y[2].x * y[0].y * y[1].z -
                    y[2].x * y[1].y * y[0].z - y[1].x * y[0].y * y[2].z - y[0].x * y[2].y * y[1].z;
  const double dM = C0 + C1 + C2 + C3;
  const bool s0 = sv_same_sign(dM, C0), s1 = sv_same_sign(dM, C1), s2 = sv_same_sign(dM, C2),
             s3 = sv_same_sign(dM, C3);
  if (s0 && s1 && s2 && s3) {
    lam.w0 = C0 / dM;
    lam.w1 = C1 / dM;
    lam.w2 = C2 / dM;
    lam.w3 = C3 / dM;
    return 15u;
  }
  double dbest = 1.e9;
  unsigned best = 0;
#pragma nounroll
  for (int j = 0; j < 4; ++j) {
    if (j == 0 ? s0 : (j == 1 ? s1 : (j == 2 ? s2 : s3))) continue;
    // the face without point j: points q0 < q1 < q2
    const int q0 = j == 0 ? 1 : 0, q1 = j <= 1 ? 2 : 1, q2 = j <= 2 ? 3 : 2;
    const d3 f0 = pick(j == 0, y[1], y[0]), f1 = pick(j <= 1, y[2], y[1]), f2 = pick(j <= 2, y[3], y[2]);
    double w0, w1, w2;
    const unsigned k = sv_tri(f0, f1, f2, w0, w1, w2);
    d3 np{0., 0., 0.};
    if (k & 1u) np += w0 * f0;
    if (k & 2u) np += w1 * f1;
    if (k & 4u) np += w2 * f2;
    const double ds = dot(np, np);
    if (ds < dbest) {
      dbest = ds;
      best = ((k & 1u) ? 1u << q0 : 0u) | ((k & 2u) ? 1u << q1 : 0u) | ((k & 4u) ? 1u << q2 : 0u);
      lam.w0 = q0 == 0 ? w0 : lam.w0;
      lam.w1 = q0 == 1 ? w0 : (q1 == 1 ? w1 : lam.w1);
      lam.w2 = q1 == 2 ? w1 : (q2 == 2 ? w2 : lam.w2);
      lam.w3 = q2 == 3 ? w2 : lam.w3;
    }
  }
  return best;
}

// the point of slot k, selected among the values
DEM_HD d3 sv_slot(const d3 (&y)[4], int k) { return pick(k == 0, y[0], pick(k == 1, y[1], pick(k == 2, y[2], y[3]))); }
// weight w into slot k of lam
DEM_HD void sv_put(L4& lam, int k, double w) {
  lam.w0 = k == 0 ? w : lam.w0;
  lam.w1 = k == 1 ? w : lam.w1;
  lam.w2 = k == 2 ? w : lam.w2;
  lam.w3 = k == 3 ? w : lam.w3;
}

// sv_subalgorithm (GJK_SV.cpp:380-410) by the number of points of the
// simplex, its points taken in slot order
DEM_HD void sv_sub(const d3 (&y)[4], unsigned& bits, L4& lam) {
  const int n = __builtin_popcount(bits);
  if (n == 1) {
    sv_put(lam, __builtin_ctz(bits), 1.);
  } else if (n == 2) {
    const int k0 = __builtin_ctz(bits), k1 = 31 - __builtin_clz(bits);
    double w0, w1;
    const unsigned k = sv_edge(sv_slot(y, k0), sv_slot(y, k1), w0, w1);
    sv_put(lam, k0, w0);
    sv_put(lam, k1, w1);
    bits = ((k & 1u) ? 1u << k0 : 0u) | ((k & 2u) ? 1u << k1 : 0u);
  } else if (n == 3) {
    const int m = __builtin_ctz(~bits & 15u);
    const int q0 = m == 0 ? 1 : 0, q1 = m <= 1 ? 2 : 1, q2 = m <= 2 ? 3 : 2;
    double w0, w1, w2;
    const unsigned k = sv_tri(sv_slot(y, q0), sv_slot(y, q1), sv_slot(y, q2), w0, w1, w2);
    sv_put(lam, q0, w0);
    sv_put(lam, q1, w1);
    sv_put(lam, q2, w2);
    bits = ((k & 1u) ? 1u << q0 : 0u) | ((k & 2u) ? 1u << q1 : 0u) | ((k & 4u) ? 1u << q2 : 0u);
  } else {
    bits = sv_tet(y, lam);
  }
}

// The search of one iteration of closest_points_GJK_SV with the new point
// in slot LAST (a compile-time index, as in gjk_iter): 1 when the point
// joins the simplex (y[LAST] set, bits updated), 0 when the loop ends, 2
// when the pair is proven farther apart than sep.
template <int LAST>
DEM_HD int sv_iter(SvSimplex& S, const ConvexGeo& A, const Tf& a2w, const ConvexGeo& B, const Tf& b2w,
                   unsigned& bits, const d3& v, double dist, double sep, int it, d3& d, d3& w, bool& accel) {
  const int go = gjk_search(A, a2w, B, b2w, A.opt, it, v, dist, sep, d, w, accel, S.p[LAST], S.q[LAST]);
  if (go != 1) return go;
  bool degen = false;
  sfor<0, 4>([&](auto I) {
    constexpr int i = decltype(I)::value;
    if ((bits & (1u << i)) && S.y[i].x == w.x && S.y[i].y == w.y && S.y[i].z == w.z) degen = true;
  });
  if (degen) return 0;
  S.y[LAST] = w;
  bits |= 1u << LAST;
  return 1;
}

// closest_points_GJK_SV (GJK_SV.cpp:428-525), tolerance and acceleration of
// A.opt; witness points in the local frames; returns the distance
DEM_HD double gjk_sv_closest(const ConvexGeo& A, const Tf& a2w, const ConvexGeo& B, const Tf& b2w, d3& pa, d3& pb,
                             int& iters, double sep, bool* separated) {
  SvSimplex S;
  sfor<0, 4>([&](auto I) {
    constexpr int i = decltype(I)::value;
    S.y[i] = S.p[i] = S.q[i] = d3{0., 0., 0.};
  });
  S.lam = L4{0., 0., 0., 0.};
  unsigned bits = 0;
  const d3 zero{0., 0., 0.};
  d3 v = tf_apply(a2w, support(A, zero)) - tf_apply(b2w, support(B, zero));
  d3 w = v, d = v;
  bool accel = A.opt.accel != 0;
  double dist = norm(v);
  int it = 0;
  while (bits < 15 && dist > kEps2 && it < 1000) {
    ++it;
    const bool plain = !accel;
    const unsigned bits0 = bits;
    const d3 v0 = v;
    int go;
    if (!(bits & 1u)) go = sv_iter<0>(S, A, a2w, B, b2w, bits, v, dist, sep, it, d, w, accel);
    else if (!(bits & 2u)) go = sv_iter<1>(S, A, a2w, B, b2w, bits, v, dist, sep, it, d, w, accel);
    else if (!(bits & 4u)) go = sv_iter<2>(S, A, a2w, B, b2w, bits, v, dist, sep, it, d, w, accel);
    else go = sv_iter<3>(S, A, a2w, B, b2w, bits, v, dist, sep, it, d, w, accel);
    if (go == 2) {
      if (separated) *separated = true;
      iters = it;
      return dist;
    }
    if (!go) break;
    // one copy of the sub-distance for every slot the point went to
    sv_sub(S.y, bits, S.lam);
    v = sv_point(S.y, bits, S.lam);
    dist = norm(v);
    // A step searched along v that drops its new point again and leaves v
    // bit for bit as it was is a fixed point: the next step sees the same
    // simplex and direction, so every step up to the cap repeats it, weights
    // included.  The loop ends here as it would at the cap (a body against
    // a long thin box meets this within ten steps).
    if (plain && bits == bits0 && same_bits(v, v0)) {
      it = 1000;
      break;
    }
  }
  pa = d3{0., 0., 0.};
  pb = d3{0., 0., 0.};
  sfor<0, 4>([&](auto I) {
    constexpr int i = decltype(I)::value;
    if (bits & (1u << i)) {
      pa += lw<i>(S.lam) * S.p[i];
      pb += lw<i>(S.lam) * S.q[i];
    }
  });
  iters = it;
  return dist;
}

// --- Bounding-volume pre-tests (OBB.cpp:144-198, OBC.cpp:226-363) -----------
// On the full transforms (RigidBodyWithCrust::isContactBVolume,
// RigidBodyWithCrust.cpp:1095-1100).

// separating-axis test of two OBBs (Gottschalk et al. 1996): the face
// normals of both and their nine cross products, with a 1e-6 margin on the
// absolute rotation
DEM_HD bool obb_overlap(const ConvexGeo& A, const ConvexGeo& B) {
  constexpr double kLowEps = 1.0e-6;
  const double a[3] = {A.obb.x, A.obb.y, A.obb.z}, b[3] = {B.obb.x, B.obb.y, B.obb.z};
  const d3 dc = B.o - A.o;
  double cen[3], ori[3][3], oa[3][3];
  for (int i = 0; i < 3; ++i) cen[i] = A.B[i] * dc.x + A.B[3 + i] * dc.y + A.B[6 + i] * dc.z;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      ori[i][j] = A.B[i] * B.B[j] + A.B[3 + i] * B.B[3 + j] + A.B[6 + i] * B.B[6 + j];
      oa[i][j] = fabs(ori[i][j]) + kLowEps;
    }
  for (int i = 0; i < 3; ++i)
    if (fabs(cen[i]) > (a[i] + b[0] * oa[i][0] + b[1] * oa[i][1] + b[2] * oa[i][2])) return false;
  for (int i = 0; i < 3; ++i)
    if (fabs(cen[0] * ori[0][i] + cen[1] * ori[1][i] + cen[2] * ori[2][i]) >
        (b[i] + a[0] * oa[0][i] + a[1] * oa[1][i] + a[2] * oa[2][i]))
      return false;
  for (int j = 0; j < 3; ++j)
    for (int i = 0; i < 3; ++i) {
      const int i1 = (i + 1) % 3, i2 = (i + 2) % 3, j1 = (j + 1) % 3, j2 = (j + 2) % 3;
      if (fabs(cen[i2] * ori[i1][j] - cen[i1] * ori[i2][j]) >
          (a[i1] * oa[i2][j] + a[i2] * oa[i1][j] + b[j1] * oa[i][j2] + b[j2] * oa[i][j1]))
        return false;
    }
  return true;
}

// OBC of a body from its OBB extents (Box/Polyhedron::computeBVolume): the
// axis along the extent closest to another one, height twice that extent,
// radius the norm of the two others; spheres: radius r, height 2r, axis x
DEM_HD void obc_of(const ConvexGeo& g, double& r, double& h, d3& e) {
  if (g.type == kShapeSphere) {
    r = g.r;
    h = 2. * g.r;
    e = d3{1., 0., 0.};
    return;
  }
  const double ex = g.obb.x, ey = g.obb.y, ez = g.obb.z;
  double xy = fabs(ex - ey);
  const double xz = fabs(ex - ez), yz = fabs(ey - ez);
  int z = xy < xz ? 2 : 1;
  xy = xy < xz ? xy : xz;
  z = xy < yz ? z : 0;
  const double ez2 = z == 0 ? ex : (z == 1 ? ey : ez);
  e = d3{z == 0 ? 1. : 0., z == 1 ? 1. : 0., z == 2 ? 1. : 0.};
  h = 2. * ez2;
  r = sqrt(ex * ex + ey * ey + ez * ez - ez2 * ez2);
}

// real roots of x^4 + b x^3 + c x^2 + d x + e: depressed quartic, the
// biquadratic case or Ferrari's resolvent cubic (OBC.cpp:143-222)
DEM_HD int quartic_roots(double b, double c, double d, double e, double (&sol)[4]) {
  int nb = 0;
  const double b2 = b * b;
  const double p = c - 3. * b2 / 8.;
  const double q = b2 * b / 8. - b * c / 2. + d;
  const double r = -3. * b2 * b2 / 256. + e - b * d / 4. + b2 * c / 16.;
  const double p2 = p * p;
  if (fabs(q) < kEps) {
    const double del = p2 / 4. - r;
    if (del < 0.) return 0;
    const double m1 = -p / 2. + sqrt(del), m2 = -p / 2. - sqrt(del);
    if (m1 > 0.) {
      sol[nb++] = sqrt(m1) - b / 4.;
      sol[nb++] = -sqrt(m1) - b / 4.;
    }
    if (m2 > 0.) {
      sol[nb++] = sqrt(m2) - b / 4.;
      sol[nb++] = -sqrt(m2) - b / 4.;
    }
    return nb;
  }
  const double u = -p2 / 36. - r / 3.;
  const double v = -p2 * p / 216. + r * p / 6. - q * q / 16.;
  const double del = u * u * u + v * v;
  double m;
  if (del < 0) {
    m = 2. * sqrt(-u) * cos(acos(v / sqrt(-u) / u) / 3.) - p / 3.;
  } else {
    m = cbrt(-v + sqrt(del));
    m = m - u / m - p / 3.;
  }
  if (m < 0.) return 0;
  const double sm = sqrt(m / 2.);
  const double f1 = -p / 2. - m / 2. - q / sm / 4.;
  const double f2 = f1 + q / sm / 2.;
  if (f1 > 0.) {
    sol[nb++] = sm + sqrt(f1) - b / 4.;
    sol[nb++] = sm - sqrt(f1) - b / 4.;
  }
  if (f2 > 0.) {
    sol[nb++] = -sm + sqrt(f2) - b / 4.;
    sol[nb++] = -sm - sqrt(f2) - b / 4.;
  }
  return nb;
}

// one ellipse step of the OBC test: the other cylinder's cap as an ellipse
// (semi-axis re) against this one's circle of radius rc; true when a point
// of the ellipse boundary lies outside the circle
DEM_HD bool obc_ellipse_apart(double x0, double yTop, double yBottom, double ez, double ey, double re, double rc) {
  if (__builtin_signbit(yTop) != __builtin_signbit(yBottom)) return false;
  const double y0 = fabs(yTop) > fabs(yBottom) ? yBottom : yTop;
  if (!(x0 * x0 + y0 * y0 / ez / ez > re * re)) return false;
  const double ey_sq = ey * ey;
  const double A = re * re * ey_sq * ey_sq;
  const double B = 2. * y0 * ez / re / ey_sq;
  const double C = x0 * x0 / A;
  const double D = y0 * y0 * ez * ez / A;
  double sint[4];
  const int nb = quartic_roots(-B, C + D - 1., B, -D, sint);
  const double sx = (0. < x0) - (x0 < 0.);
  for (int i = 0; i < 4; ++i)
    if (i < nb && __builtin_signbit(sint[i]) != __builtin_signbit(B)) {
      const double cost = -sx * sqrt(1. - sint[i] * sint[i]);
      const double ptX = x0 + re * cost;
      const double ptY = y0 + re * ez * sint[i];
      if (ptX * ptX + ptY * ptY > rc * rc) return true;
    }
  return false;
}

DEM_HD d3 basis_times(const double* B, d3 v) {
  return d3{B[0] * v.x + B[1] * v.y + B[2] * v.z, B[3] * v.x + B[4] * v.y + B[5] * v.z,
            B[6] * v.x + B[7] * v.y + B[8] * v.z};
}
DEM_HD d3 unit_of(d3 v) { return v / norm(v); }

// isContactBVolume(OBC, OBC) (OBC.cpp:226-363): shortest distance of the
// axes, projections of the two rectangles, then ellipse against circle
DEM_HD bool obc_overlap(const ConvexGeo& A, const ConvexGeo& B) {
  double r1, hh1, r2, hh2;
  d3 e1, e2l;
  obc_of(A, r1, hh1, e1);
  obc_of(B, r2, hh2, e2l);
  const double h1 = hh1 / 2., h2 = hh2 / 2.;
  const d3 eZ = basis_times(A.B, e1);
  const d3 e2 = basis_times(B.B, e2l);
  const d3 xPt = B.o - A.o;
  const d3 eX = unit_of(cross(e2, eZ));
  const d3 eY = unit_of(cross(eZ, eX));
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
    if (obc_ellipse_apart(x, y + h2 * ey, y - h2 * ey, ez, ey, r2, r1)) return false;
    const d3 eY2 = unit_of(cross(eX, e2));
    const double y2 = -dot(eY2, xPt);
    const double ey2 = dot(eY2, eZ);
    if (obc_ellipse_apart(-x, y2 + h1 * ey2, y2 - h1 * ey2, ez, ey2, r1, r2)) return false;
  }
  return true;
}

DEM_HD bool bv_overlap(const ConvexGeo& A, const ConvexGeo& B) {
  return A.opt.bv == 1 ? obb_overlap(A, B) : obc_overlap(A, B);
}

// closest_points with Johnson's sub-distance (Convex.cpp:536-627), the
// tolerance and acceleration of A.opt; witness points in the local frames;
// returns the distance (sep >= 0: *separated set when the distance is
// proven larger than sep; the witness points are then not computed)
DEM_HD double gjk_johnson_closest(const ConvexGeo& A, const Tf& a2w, const ConvexGeo& B, const Tf& b2w, d3& pa,
                                  d3& pb, int& iters, double sep, bool* separated) {
  Simplex S;
  for (int i = 0; i < 16; ++i)
    for (int j = 0; j < 4; ++j) S.det[i][j] = 0.;
  for (int i = 0; i < 10; ++i) S.dp[i] = 0.;
  unsigned bits = 0, all_bits = 0;
  d3 zero{0., 0., 0.};
  d3 v = tf_apply(a2w, support(A, zero)) - tf_apply(b2w, support(B, zero));
  d3 w = v, d = v;
  bool accel = A.opt.accel != 0;
  double dist = norm(v);
  int it = 0;
  while (bits < 15 && dist > kEps2 && it < 1000) {
    ++it;
    // the first free slot
    int go;
    if (!(bits & 1u)) go = gjk_iter<0>(S, A, a2w, B, b2w, bits, all_bits, v, dist, sep, it, d, w, accel);
    else if (!(bits & 2u)) go = gjk_iter<1>(S, A, a2w, B, b2w, bits, all_bits, v, dist, sep, it, d, w, accel);
    else if (!(bits & 4u)) go = gjk_iter<2>(S, A, a2w, B, b2w, bits, all_bits, v, dist, sep, it, d, w, accel);
    else go = gjk_iter<3>(S, A, a2w, B, b2w, bits, all_bits, v, dist, sep, it, d, w, accel);
    if (go == 2) {
      if (separated) *separated = true;
      iters = it;
      return dist;
    }
    if (!go) break;
  }
  j_points(S, bits, pa, pb);
  iters = it;
  return dist;
}

// The GJK distance the system's options choose: the signed-volume
// sub-distance (GJK_SV) or Johnson's; johnson: Johnson's whatever the
// options (the ContactError recovery, RigidBodyWithCrust.cpp:545).
DEM_HD double gjk_closest(const ConvexGeo& A, const Tf& a2w, const ConvexGeo& B, const Tf& b2w, d3& pa, d3& pb,
                          int& iters, double sep = -1., bool* separated = nullptr, bool johnson = false) {
  if (A.opt.sv && !johnson) return gjk_sv_closest(A, a2w, B, b2w, pa, pb, iters, sep, separated);
  return gjk_johnson_closest(A, a2w, B, b2w, pa, pb, iters, sep, separated);
}

// General ClosestPoint path.  Returns 1 with the contact features (the
// caller tests d < 0), 0 when the circumscribed spheres (or, with the
// option, the bounding volumes) do not meet, -1 on ContactError (the
// crust-shrunk bodies intersect).
DEM_HD int closest_gjk(const ConvexGeo& A, const ConvexGeo& B, d3& P, d3& ov, double& d, int* iters = nullptr) {
  d3 gcagcb = A.o - B.o;
  if (!(norm(gcagcb) < A.circR + B.circR)) return 0;
  if (A.opt.bv && !bv_overlap(A, B)) return 0;
  Tf a2w = tf_scaled(A, A.scaling), b2w = tf_scaled(B, B.scaling);
  d3 pa, pb;
  int it;
  double distance = gjk_closest(A, a2w, B, b2w, pa, pb, it);
  if (iters) *iters = it;
  if (distance < kEps) return -1;
  pa = tf_apply(a2w, pa);
  pb = tf_apply(b2w, pb);
  d3 ba = pa - pb;
  P = pa / 2.0 + pb / 2.0;
  d3 o = ba / distance;
  round_small(o);
  o *= A.crust + B.crust - distance;
  ov = o;
  d = distance - (A.crust + B.crust);
  return 1;
}

// ClosestPoint_ErreurHandling(factor 10): false if still intersecting
DEM_HD bool recover_gjk(const ConvexGeo& A, const ConvexGeo& B, d3& P, d3& ov, double& d, int* iters = nullptr) {
  d3 sa, sb;
  double* as[3] = {&sa.x, &sa.y, &sa.z};
  double* bs[3] = {&sb.x, &sb.y, &sb.z};
  const double ae[3] = {A.bext.x, A.bext.y, A.bext.z}, be[3] = {B.bext.x, B.bext.y, B.bext.z};
  for (int a = 0; a < 3; ++a) {
    double x = ae[a] < kEps ? 1. : (ae[a] - 10. * A.crust) / ae[a];
    *as[a] = x > 0.5 ? x : 0.5;
    double y = be[a] < kEps ? 1. : (be[a] - 10. * B.crust) / be[a];
    *bs[a] = y > 0.5 ? y : 0.5;
  }
  Tf a2w = tf_scaled(A, sa), b2w = tf_scaled(B, sb);
  d3 pa, pb;
  int it;
  double distance = gjk_closest(A, a2w, B, b2w, pa, pb, it, -1., nullptr, true);
  if (iters) *iters = it;
  if (distance < kEps) return false;
  pa = tf_apply(a2w, pa);
  pb = tf_apply(b2w, pb);
  d3 ba = pa - pb;
  ba /= norm(ba);
  P = pa / 2.0 + pb / 2.0;
  double imposed = 1. * (A.crust + B.crust);
  ov = imposed * ba;
  d = -imposed;
  return true;
}

// One attempt of the general ClosestPoint path: attempt 0 with the
// crust-shrunk transforms (closest_gjk), attempt 1 with the bodies shrunk by
// ten crusts (recover_gjk, the ContactError recovery).  Returns 1 with the
// contact features (d may be >= 0), 2 with the recovered features, 0 when
// the circumscribed spheres do not meet (attempt 0), -1 when the shrunk
// bodies intersect (attempt 0: go on with attempt 1; attempt 1:
// ContactError).  A caller that evaluates every pair through one call of
// this function holds a single copy of the GJK state in registers.
// early: attempt 0 stops as soon as the pair is proven farther apart than
// the crusts (no contact; the iteration count is then not the full run's,
// so a caller collecting statistics passes false).
DEM_HD int gjk_attempt(const ConvexGeo& A, const ConvexGeo& B, int attempt, d3& P, d3& ov, double& d,
                       int& iters, bool early = false) {
  d3 sa = A.scaling, sb = B.scaling;
  if (attempt == 0) {
    d3 gcagcb = A.o - B.o;
    if (!(norm(gcagcb) < A.circR + B.circR)) return 0;
    if (A.opt.bv && !bv_overlap(A, B)) return 0;
  } else {
    double* as[3] = {&sa.x, &sa.y, &sa.z};
    double* bs[3] = {&sb.x, &sb.y, &sb.z};
    const double ae[3] = {A.bext.x, A.bext.y, A.bext.z}, be[3] = {B.bext.x, B.bext.y, B.bext.z};
    for (int a = 0; a < 3; ++a) {
      double x = ae[a] < kEps ? 1. : (ae[a] - 10. * A.crust) / ae[a];
      *as[a] = x > 0.5 ? x : 0.5;
      double y = be[a] < kEps ? 1. : (be[a] - 10. * B.crust) / be[a];
      *bs[a] = y > 0.5 ? y : 0.5;
    }
  }
  const double crusts = A.crust + B.crust;
  Tf a2w = tf_scaled(A, sa), b2w = tf_scaled(B, sb);
  d3 pa, pb;
  bool separated = false;
  double distance =
      gjk_closest(A, a2w, B, b2w, pa, pb, iters, early && attempt == 0 ? crusts : -1., &separated, attempt != 0);
  if (separated) return 0;
  if (distance < kEps) return -1;
  pa = tf_apply(a2w, pa);
  pb = tf_apply(b2w, pb);
  d3 ba = pa - pb;
  if (attempt == 0) {
    P = pa / 2.0 + pb / 2.0;
    d3 o = ba / distance;
    round_small(o);
    o *= crusts - distance;
    ov = o;
    d = distance - crusts;
    return 1;
  }
  ba /= norm(ba);
  P = pa / 2.0 + pb / 2.0;
  double imposed = 1. * crusts;
  ov = imposed * ba;
  d = -imposed;
  return 2;
}

// --- kinematics of non-spherical bodies --------------------------------------
DEM_HD d3 ang_acc3(d3 tbf, double cf, d3 I, d3 om) {
  return d3{tbf.x / (cf * I.x) + om.y * om.z * (I.y - I.z) / I.x, tbf.y / (cf * I.y) + om.z * om.x * (I.z - I.x) / I.y,
            tbf.z / (cf * I.z) + om.x * om.y * (I.x - I.y) / I.z};
}
DEM_HD void ssprk3_3(d3& vrot, d3 tbf, double cf, d3 I, double dt) {
  d3 K1 = ang_acc3(tbf, cf, I, vrot);
  K1 *= dt;
  d3 K2 = ang_acc3(tbf, cf, I, vrot + K1);
  K2 *= dt;
  d3 K3 = ang_acc3(tbf, cf, I, vrot + 0.25 * (K1 + K2));
  K3 *= dt;
  vrot += (K1 + K2 + 4. * K3) / 6.;
}
// basis <- R(q) basis
DEM_HD void compose_left(double* B, quat q) {
  const double qx = q.v.x, qy = q.v.y, qz = q.v.z, qw = q.w;
  for (int i = 0; i < 3; ++i) {
    double px = qy * B[6 + i] - qz * B[3 + i] + qw * B[i];
    double py = qz * B[i] - qx * B[6 + i] + qw * B[3 + i];
    double pz = qx * B[3 + i] - qy * B[i] + qw * B[6 + i];
    double pw = -qx * B[i] - qy * B[3 + i] - qz * B[6 + i];
    B[i] = qy * pz - qz * py - pw * qx + qw * px;
    B[3 + i] = qz * px - qx * pz - pw * qy + qw * py;
    B[6 + i] = qx * py - qy * px - pw * qz + qw * pz;
  }
}

}  // namespace dem
