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
template <int LAST>
DEM_HD int gjk_iter(Simplex& S, const ConvexGeo& A, const Tf& a2w, const ConvexGeo& B, const Tf& b2w,
                    unsigned& bits, unsigned& all_bits, d3& v, double& dist, double sep) {
  constexpr unsigned last_bit = 1u << LAST;
  const double relError = kEps, absError = 1.e-4 * relError;
  d3 d = v;
  S.p[LAST] = support(A, tf_vec_basis(-d, a2w));
  S.q[LAST] = support(B, tf_vec_basis(d, b2w));
  d3 w = tf_apply(a2w, S.p[LAST]) - tf_apply(b2w, S.q[LAST]);
  if (sep >= 0. && dot(v, w) > sep * dist + 1.e-12 * dist * (fabs(w.x) + fabs(w.y) + fabs(w.z))) return 2;
  double mu = dist - dot(v, w) / dist;
  if (mu < dist * relError || mu < absError) return 0;
  if (j_degenerate(S, all_bits, w)) return 0;
  S.y[LAST] = w;
  all_bits = bits | last_bit;
  if (!j_closest<LAST>(S, bits, all_bits, v)) return 0;
  dist = norm(v);
  return 1;
}

// closest_points with relative tolerance 1e-10 (Convex.cpp:536-627);
// witness points in the local frames; returns the distance
// (sep >= 0: *separated set when the distance is proven larger than sep;
// the witness points are then not computed)
DEM_HD double gjk_closest(const ConvexGeo& A, const Tf& a2w, const ConvexGeo& B, const Tf& b2w, d3& pa, d3& pb,
                          int& iters, double sep = -1., bool* separated = nullptr) {
  Simplex S;
  for (int i = 0; i < 16; ++i)
    for (int j = 0; j < 4; ++j) S.det[i][j] = 0.;
  for (int i = 0; i < 10; ++i) S.dp[i] = 0.;
  unsigned bits = 0, all_bits = 0;
  d3 zero{0., 0., 0.};
  d3 v = tf_apply(a2w, support(A, zero)) - tf_apply(b2w, support(B, zero));
  double dist = norm(v);
  int it = 0;
  while (bits < 15 && dist > kEps2 && it < 1000) {
    ++it;
    // the first free slot
    int go;
    if (!(bits & 1u)) go = gjk_iter<0>(S, A, a2w, B, b2w, bits, all_bits, v, dist, sep);
    else if (!(bits & 2u)) go = gjk_iter<1>(S, A, a2w, B, b2w, bits, all_bits, v, dist, sep);
    else if (!(bits & 4u)) go = gjk_iter<2>(S, A, a2w, B, b2w, bits, all_bits, v, dist, sep);
    else go = gjk_iter<3>(S, A, a2w, B, b2w, bits, all_bits, v, dist, sep);
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

// General ClosestPoint path.  Returns 1 with the contact features (the
// caller tests d < 0), 0 when the circumscribed spheres do not meet, -1 on
// ContactError (the crust-shrunk bodies intersect).
DEM_HD int closest_gjk(const ConvexGeo& A, const ConvexGeo& B, d3& P, d3& ov, double& d, int* iters = nullptr) {
  d3 gcagcb = A.o - B.o;
  if (!(norm(gcagcb) < A.circR + B.circR)) return 0;
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
  double distance = gjk_closest(A, a2w, B, b2w, pa, pb, it);
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
  double distance = gjk_closest(A, a2w, B, b2w, pa, pb, iters, early && attempt == 0 ? crusts : -1., &separated);
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
