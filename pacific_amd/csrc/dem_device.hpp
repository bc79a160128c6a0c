// Device functions of the DEM step shared by the single-system kernels
// (dem_kernels.hip) and the batched persistent kernel (dem_batch.hip):
// narrow phase, Hooke-dashpot-Coulomb force model, integrators, binning.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

#include "dem_convex.hpp"
#include "dem_math.hpp"
#include "dem_types.hpp"

namespace dem {

// ---------------------------------------------------------------------------
// FP32 helpers for the mixed force path
struct f3 {
  float x, y, z;
};
DEM_HD f3 f3of(d3 a) { return f3{(float)a.x, (float)a.y, (float)a.z}; }
DEM_HD f3 operator+(f3 a, f3 b) { return f3{a.x + b.x, a.y + b.y, a.z + b.z}; }
DEM_HD f3 operator-(f3 a, f3 b) { return f3{a.x - b.x, a.y - b.y, a.z - b.z}; }
DEM_HD f3 operator-(f3 a) { return f3{-a.x, -a.y, -a.z}; }
DEM_HD f3 operator*(f3 a, float d) { return f3{a.x * d, a.y * d, a.z * d}; }
DEM_HD f3 operator*(float d, f3 a) { return f3{a.x * d, a.y * d, a.z * d}; }
DEM_HD f3& operator+=(f3& a, f3 b) { a.x += b.x; a.y += b.y; a.z += b.z; return a; }
DEM_HD float dotf(f3 a, f3 b) { return fmaf(a.z, b.z, fmaf(a.y, b.y, a.x * b.x)); }
DEM_HD f3 crossf(f3 a, f3 b) {
  return f3{fmaf(a.y, b.z, -a.z * b.y), fmaf(a.z, b.x, -a.x * b.z), fmaf(a.x, b.y, -a.y * b.x)};
}
DEM_HD float normf(f3 a) { return sqrtf(dotf(a, a)); }

// ---------------------------------------------------------------------------
// Workgroup -> tile mapping.  Consecutive workgroups are dispatched
// round-robin over the 8 XCDs; remapping gives each XCD one contiguous range
// of sorted particles, so neighbouring cells are fetched into the same L2.
__device__ __forceinline__ int xcd_tile(int b, int nb) {
  int xcd = b & 7, idx = b >> 3;
  int per = nb >> 3, rem = nb & 7;
  return xcd * per + (xcd < rem ? xcd : rem) + idx;
}

__device__ __forceinline__ void raise_error(DevStatus* st, int code, int id0, int id1) {
  if (atomicCAS(&st->code, 0, code) == 0) {
    st->id0 = id0;
    st->id1 = id1;
  }
}

// Cell::GetCell (Cell.cpp:172-184) on one axis, in reference indexing
__device__ __forceinline__ int ref_cell(double x, const Grid& g, int a) {
  return int(floor((x - g.lo[a]) / g.edge[a]));
}

// ---------------------------------------------------------------------------
// Narrow phase (FP64)

// ClosestPointSPHERESPHERE (RigidBodyWithCrust.cpp:596-640) with A = self.
// The caller has already computed AB = xB - xA and nab = Norm(AB).
// Returns 1 contact, 0 none, -1 ContactError (overlap >= crust sum).
DEM_HD int sphere_sphere(d3 xA, double rA, double cA, d3 AB, double nab, double rB, double cB, d3& P, d3& ov,
                         double& d) {
  double distance = nab - (rA + rB);
  if (distance > 0.) return 0;
  if (-distance >= cA + cB) return -1;
  P = xA + ((rA + 0.5 * distance) * AB) / nab;
  ov = distance * (AB / nab);
  d = distance;
  return 1;
}

// Box corner n in the numbering of Box::setCornersFaces (Box.cpp:92-110)
DEM_HD d3 corner_of(const double* e, int n) {
  const bool px = n == 1 || n == 2 || n == 5 || n == 6, py = n >= 4, pz = n == 2 || n == 3 || n == 6 || n == 7;
  return d3{px ? e[0] : -e[0], py ? e[1] : -e[1], pz ? e[2] : -e[2]};
}
DEM_HD double comp(d3 a, int i) { return i == 0 ? a.x : (i == 1 ? a.y : a.z); }
DEM_HD void set_comp(d3& a, int i, double v) {
  if (i == 0) a.x = v; else if (i == 1) a.y = v; else a.z = v;
}

// Box::ContactCornerSPHERE (Box.cpp:725-744)
DEM_HD void box_corner(const double* e, d3 C, double R, int n, double& overlap, d3& cp) {
  d3 dist = C - corner_of(e, n);
  double nd = norm(dist);
  if (nd < R) {
    overlap = nd - R;
    cp = C - (1. + 0.5 * overlap / nd) * dist;
  }
}
// Box::ContactEdgeSPHERE (Box.cpp:750-769)
DEM_HD void box_edge(const double* e, d3 C, double R, int n, int dir, double& overlap, d3& cp) {
  d3 proj = C;
  d3 cn = corner_of(e, n);
  set_comp(proj, dir, comp(cn, dir));
  d3 dist = proj - cn;
  double nd = norm(dist);
  if (nd < R) {
    overlap = nd - R;
    cp = C - (1. + 0.5 * overlap / nd) * dist;
  }
}
// Box::IntersectionPointSPHERE (Box.cpp:466-716); false if the centre lies in
// the box (a ContactError in the reference)
DEM_HD bool box_sphere(const double* e, d3 C, double R, double& overlap, d3& cp) {
  const double ex = e[0], ey = e[1], ez = e[2];
  const double gx = C.x, gy = C.y, gz = C.z;
  overlap = 1.;
  cp = d3{0., 0., 0.};
  if (gx > ex) {
    if (gy > ey) {
      if (gz > ez) box_corner(e, C, R, 6, overlap, cp);
      else if (gz < -ez) box_corner(e, C, R, 5, overlap, cp);
      else box_edge(e, C, R, 5, 2, overlap, cp);
    } else if (gy < -ey) {
      if (gz > ez) box_corner(e, C, R, 2, overlap, cp);
      else if (gz < -ez) box_corner(e, C, R, 1, overlap, cp);
      else box_edge(e, C, R, 1, 2, overlap, cp);
    } else {
      if (gz > ez) box_edge(e, C, R, 6, 1, overlap, cp);
      else if (gz < -ez) box_edge(e, C, R, 1, 1, overlap, cp);
      else {
        double nd = gx - ex;
        if (nd < R) { overlap = nd - R; cp = d3{ex + 0.5 * overlap, C.y, C.z}; }
      }
    }
  } else if (gx < -ex) {
    if (gy > ey) {
      if (gz > ez) box_corner(e, C, R, 7, overlap, cp);
      else if (gz < -ez) box_corner(e, C, R, 4, overlap, cp);
      else box_edge(e, C, R, 4, 2, overlap, cp);
    } else if (gy < -ey) {
      if (gz > ez) box_corner(e, C, R, 3, overlap, cp);
      else if (gz < -ez) box_corner(e, C, R, 0, overlap, cp);
      else box_edge(e, C, R, 0, 2, overlap, cp);
    } else {
      if (gz > ez) box_edge(e, C, R, 3, 1, overlap, cp);
      else if (gz < -ez) box_edge(e, C, R, 0, 1, overlap, cp);
      else {
        double nd = -ex - gx;
        if (nd < R) { overlap = nd - R; cp = d3{-ex - 0.5 * overlap, C.y, C.z}; }
      }
    }
  } else {
    if (gy > ey) {
      if (gz > ez) box_edge(e, C, R, 6, 0, overlap, cp);
      else if (gz < -ez) box_edge(e, C, R, 4, 0, overlap, cp);
      else {
        double nd = gy - ey;
        if (nd < R) { overlap = nd - R; cp = d3{C.x, ey + 0.5 * overlap, C.z}; }
      }
    } else if (gy < -ey) {
      if (gz > ez) box_edge(e, C, R, 3, 0, overlap, cp);
      else if (gz < -ez) box_edge(e, C, R, 0, 0, overlap, cp);
      else {
        double nd = -ey - gy;
        if (nd < R) { overlap = nd - R; cp = d3{C.x, -ey - 0.5 * overlap, C.z}; }
      }
    } else {
      if (gz > ez) {
        double nd = gz - ez;
        if (nd < R) { overlap = nd - R; cp = d3{C.x, C.y, ez + 0.5 * overlap}; }
      } else if (gz < -ez) {
        double nd = -ez - gz;
        if (nd < R) { overlap = nd - R; cp = d3{C.x, C.y, -ez - 0.5 * overlap}; }
      } else {
        return false;
      }
    }
  }
  return true;
}

// Transform::operator() (Transform.cpp:310-316)
DEM_HD d3 apply_tf(const double* B, const double* o, d3 p) {
  return d3{dot(d3{B[0], B[1], B[2]}, p) + o[0], dot(d3{B[3], B[4], B[5]}, p) + o[1],
            dot(d3{B[6], B[7], B[8]}, p) + o[2]};
}

// ClosestPointSPHEREBOX with the sphere as body A (RigidBodyWithCrust.cpp:647-693)
DEM_HD int sphere_box(const BoxObstacle& o, d3 xA, double rA, double cA, d3& P, d3& ov, double& d) {
  double overlap = 0.;
  d3 local = apply_tf(o.Binv, o.invorigin, xA);
  d3 cpl;
  if (!box_sphere(o.extent, local, rA, overlap, cpl)) return -1;
  if (!(overlap < 0.)) return 0;
  if (-overlap >= cA + o.crust) return -1;
  P = apply_tf(o.B, o.centre, cpl);
  d3 AB = P - xA;
  ov = (overlap / norm(AB)) * AB;
  d = overlap;
  return 1;
}

// ContactError recovery, RigidBodyWithCrust::ClosestPoint_ErreurHandling
// (RigidBodyWithCrust.cpp:534-590): distance between the shapes shrunk by ten
// crust thicknesses (scale capped at 0.5, :770-798), closed form for spheres
// and boxes.  Returns false when the shrunk shapes still overlap.
DEM_HD double shrink_scale(double extent, double crust) {
  double sc = extent < kEps ? 1. : (extent - 10. * crust) / extent;
  return sc > 0.5 ? sc : 0.5;
}
DEM_HD bool recover_ss(d3 xA, double rA, double cA, d3 xB, double rB, double cB, d3& P, d3& ov, double& d) {
  double ra = rA * shrink_scale(rA, cA);
  double rb = rB * shrink_scale(rB, cB);
  d3 AB = xB - xA;
  double nab = norm(AB);
  double distance = nab - ra - rb;
  if (distance < kEps) return false;
  d3 n = AB / nab;
  d3 pA = xA + ra * n, pB = xB - rb * n;
  d3 ba = pA - pB;
  ba /= norm(ba);
  P = pA / 2.0 + pB / 2.0;
  double imposed = 1. * (cA + cB);
  ov = imposed * ba;
  d = -imposed;
  return true;
}
DEM_HD bool recover_sb(const BoxObstacle& o, d3 xA, double rA, double cA, d3& P, d3& ov, double& d) {
  double ra = rA * shrink_scale(rA, cA);
  double es[3];
  for (int a = 0; a < 3; ++a) es[a] = o.extent[a] * shrink_scale(o.extent[a], o.crust);
  d3 local = apply_tf(o.Binv, o.invorigin, xA);
  d3 q{fmin(fmax(local.x, -es[0]), es[0]), fmin(fmax(local.y, -es[1]), es[1]), fmin(fmax(local.z, -es[2]), es[2])};
  d3 dl = local - q;
  double nd = norm(dl);
  double distance = nd - ra;
  if (distance < kEps) return false;
  d3 pAl = local - (ra / nd) * dl;
  d3 pA = apply_tf(o.B, o.centre, pAl), pB = apply_tf(o.B, o.centre, q);
  d3 ba = pA - pB;
  ba /= norm(ba);
  P = pA / 2.0 + pB / 2.0;
  double imposed = 1. * (cA + o.crust);
  ov = imposed * ba;
  d = -imposed;
  return true;
}

// ClosestPointSPHEREBOX against a box particle: the sphere is body A
// (RigidBodyWithCrust.cpp:661-693) or the box is (:694-716).  The box's
// inverse transform is its transposed basis (Transform.cpp:71-77).
DEM_HD int sphere_box_particle(const ConvexGeo& bx, const ConvexGeo& sp, bool sphere_first, d3& P, d3& ov,
                               double& d) {
  double Binv[9], io[3];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) Binv[3 * i + j] = bx.B[3 * j + i];
  for (int i = 0; i < 3; ++i) io[i] = dot(-d3{Binv[3 * i], Binv[3 * i + 1], Binv[3 * i + 2]}, bx.o);
  const double e[3] = {bx.e.x, bx.e.y, bx.e.z};
  const double oc[3] = {bx.o.x, bx.o.y, bx.o.z};
  double overlap = 0.;
  d3 local = apply_tf(Binv, io, sp.o);
  d3 cpl;
  if (!box_sphere(e, local, sp.circR, overlap, cpl)) return -1;
  if (!(overlap < 0.)) return 0;
  if (-overlap >= sp.crust + bx.crust) return -1;
  P = apply_tf(bx.B, oc, cpl);
  d3 AB = sphere_first ? P - sp.o : sp.o - P;
  ov = (overlap / norm(AB)) * AB;
  d = overlap;
  return 1;
}

// RigidBodyWithCrust::ClosestPoint dispatch for a pair with at least one
// non-spherical body.  1: features set (contact if d < 0); 0: no contact;
// -1: ContactError.
// iters: iterations the reference's PointContact reports (GJK's count; 1
// for the closed forms' contact points, RigidBodyWithCrust.cpp:635, 689, 715)
DEM_HD int closest_convex(const ConvexGeo& A, const ConvexGeo& B, d3& P, d3& ov, double& d, int* iters = nullptr) {
  int r;
  if (A.type == kShapeSphere && B.type == kShapeBox) r = sphere_box_particle(B, A, true, P, ov, d);
  else if (A.type == kShapeBox && B.type == kShapeSphere) r = sphere_box_particle(A, B, false, P, ov, d);
  else return closest_gjk(A, B, P, ov, d, iters);
  if (iters) *iters = 1;
  return r;
}

// ---------------------------------------------------------------------------
// Hooke force model, FP64 (HookeContactForceModel.cpp:44-114), p0 = self.
// Returns the force on p0 and the rolling moment on p0.
// Returns true when the Coulomb cap is active.
DEM_HD bool hooke64(const HookeModel& m, d3 x0, d3 v0, d3 w0, double m0, double req0, d3 x1, d3 v1, d3 w1,
                    double m1, double req1, d3 P, d3 pen, d3& f, d3& dM) {
  d3 normal = pen;
  normal /= norm(normal);
  round_small(normal);
  d3 tmpV = (v0 + cross(w0, P - x0)) - (v1 + cross(w1, P - x1));
  d3 v_n = normal * dot(tmpV, normal);
  d3 v_t = tmpV - v_n;
  d3 delFN = m.kn * pen;
  double avmass = 1. / (1. / m0 + 1. / m1);
  double gamman = -2. * m.beta * sqrt(avmass * m.kn);
  delFN -= gamman * v_n;
  double normFN = norm(delFN);
  double etat = m.etat == -1. ? -m.beta * sqrt(m.kn / avmass) : m.etat;
  double gammat = 2. * etat * avmass;
  d3 delFT = -gammat * v_t;
  double fn = m.muc * normFN;
  double ft = norm(delFT);
  if (fn < ft) {
    double nvt = norm(v_t);
    d3 tangent{0., 0., 0.};
    if (nvt > kEps) tangent = -v_t / nvt;
    delFT = tangent * fn;
  }
  dM = d3{0., 0., 0.};
  if (m.kr != 0.) {
    d3 wrel = w0 - w1;
    double nwrel = norm(wrel);
    double nwtrel = norm(cross(w0, x0 - P) - cross(w1, x1 - P));
    if (nwrel > kEps) {
      double Req = 1. / (1. / req0 + 1. / req1);
      dM = -(m.kr * Req * normFN * nwtrel / nwrel) * wrel;
    }
  }
  double coef = 1.;
  f = coef * (delFN + delFT);
  return fn < ft;
}

// Quaternion::setRotFromTwoVectors (Quaternion.cpp:263-286): the rotation
// taking u to v, (x, y, z, w) normalised.
struct RotQ {
  d3 v;
  double w;
};
DEM_HD RotQ rot_from_two(const d3& u, const d3& v) {
  double nn = sqrt(dot(u, u) * dot(v, v));
  double re = nn + dot(u, v);
  d3 vect;
  if (re < 1.e-6 * nn) {
    re = 0.;
    vect = fabs(u.x) > fabs(u.z) ? d3{-u.y, u.x, 0.} : d3{0., -u.z, u.y};
  } else {
    vect = cross(u, v);
  }
  const double d = 1. / sqrt(vect.x * vect.x + vect.y * vect.y + vect.z * vect.z + re * re);
  return RotQ{d * vect, d * re};
}
// Quaternion::rotateVector (Quaternion.cpp:293-298)
DEM_HD d3 rotate_vec(const RotQ& q, const d3& x) {
  const double nv = norm(q.v);
  return ((q.w * q.w - nv * nv) * x + (2. * dot(x, q.v)) * q.v) + (2. * q.w) * cross(q.v, x);
}

// HookeMemoryContactForceModel (HookeMemoryContactForceModel.cpp:54-193),
// p0 = self: normal spring-dashpot, tangential spring kt with memory kd
// rotated into the current contact plane and capped at Coulomb, rolling
// spring torque sr with memory (mur != 0).  existed: the pair was in contact
// at the last evaluation (kd, pn, sr hold its memory; zero otherwise); they
// are updated in place.  Returns the force on p0 and its rolling torque;
// true when the Coulomb cap applied.
DEM_HD bool hooke_memory64(const HookeModel& m, d3 x0, d3 v0, d3 w0, double m0, double req0, d3 x1, d3 v1, d3 w1,
                           double m1, double req1, d3 P, d3 pen, double dt, bool existed, d3& kd, d3& pn, d3& sr,
                           d3& f, d3& dM) {
  d3 normal = pen;
  normal /= norm(normal);
  round_small(normal);
  d3 tmpV = (v0 + cross(w0, P - x0)) - (v1 + cross(w1, P - x1));
  d3 v_n = normal * dot(tmpV, normal);
  d3 v_t = tmpV - v_n;
  d3 delFN = m.kn * pen;
  double avmass = 1. / (1. / m0 + 1. / m1);
  double gamman = -2. * m.beta * sqrt(avmass * m.kn);
  delFN -= gamman * v_n;
  double normFN = norm(delFN);
  double etat = m.etat == -1. ? -m.beta * sqrt(m.kn / avmass) : m.etat;
  double gammat = 2. * etat * avmass;
  d3 viscousFT = -gammat * v_t;
  RotQ q{d3{0., 0., 0.}, 0.};
  if (existed) {
    q = rot_from_two(pn, normal);
    kd = rotate_vec(q, kd);
  }
  kd += (dt * m.kt) * v_t;
  pn = normal;
  // computeTangentialVector (:54-65)
  d3 tij{0., 0., 0.};
  if (norm(v_t) > 0.001) {
    tij = -v_t / norm(v_t);
  } else {
    const d3 nt = gammat * v_t, a = kd + nt;
    if ((norm(kd) > m.epst || norm(nt) > m.epst) && norm(a) > m.epst) tij = -a / norm(a);
  }
  d3 delFT;
  const bool slip = !(norm(-kd + viscousFT) <= m.muc * normFN);
  if (!slip) {
    delFT = norm(-kd + viscousFT) * tij;
  } else {
    if (m.kt != 0.) kd = ((-m.muc) * normFN * tij + viscousFT) / m.kt;
    delFT = (m.muc * normFN) * tij;
  }
  dM = d3{0., 0., 0.};
  if (m.mur != 0.) {
    double Req = 1. / (1. / req0 + 1. / req1);
    double kr = 3. * m.kn * m.mur * m.mur * Req * Req;
    double etar = 3. * gamman * m.mur * m.mur * Req * Req;
    double max_normMk = m.mur * Req * normFN;
    d3 wrel = w0 - w1;
    if (existed) sr = rotate_vec(q, sr);
    sr -= (kr * dt) * wrel;
    double normMk = norm(sr);
    if (normMk > max_normMk) sr *= max_normMk / normMk;
    dM = sr - (m.etarpf * etar) * wrel;
  }
  double coef = 1.;
  f = coef * (delFN + delFT);
  return slip;
}

// Velocity of body 0 relative to body 1 at the contact point P
// (Particle.cpp:661-665), in FP64: the translational velocities of two
// bodies in contact are close, and their difference in FP32 would lose the
// relative velocity the dashpot acts on.
DEM_HD d3 rel_velocity(d3 x0, d3 v0, d3 w0, d3 x1, d3 v1, d3 w1, d3 P) {
  return (v0 + cross(w0, P - x0)) - (v1 + cross(w1, P - x1));
}

// Mixed precision: the force arithmetic in FP32 on lever arms l = P - x and
// the FP64 relative velocity tmpV at the contact point.
DEM_HD bool hooke32(const HookeModel& m, f3 l0, f3 w0, float m0, float req0, f3 l1, f3 w1, float m1, float req1,
                    f3 tmpV, f3 pen, f3& f, f3& dM) {
  const float kn = (float)m.kn, beta = (float)m.beta;
  float npen = normf(pen);
  f3 normal = pen * (1.f / npen);
  normal.x = fabsf(normal.x) < 1e-10f ? 0.f : normal.x;
  normal.y = fabsf(normal.y) < 1e-10f ? 0.f : normal.y;
  normal.z = fabsf(normal.z) < 1e-10f ? 0.f : normal.z;
  f3 v_n = normal * dotf(tmpV, normal);
  f3 v_t = tmpV - v_n;
  float avmass = 1.f / (1.f / m0 + 1.f / m1);
  float gamman = -2.f * beta * sqrtf(avmass * kn);
  f3 delFN = kn * pen - gamman * v_n;
  float normFN = normf(delFN);
  float etat = m.etat == -1. ? -beta * sqrtf(kn / avmass) : (float)m.etat;
  float gammat = 2.f * etat * avmass;
  f3 delFT = -gammat * v_t;
  float fn = (float)m.muc * normFN;
  float ft = normf(delFT);
  if (fn < ft) {
    float nvt = normf(v_t);
    f3 tangent{0.f, 0.f, 0.f};
    if (nvt > 1e-10f) tangent = v_t * (-1.f / nvt);
    delFT = tangent * fn;
  }
  dM = f3{0.f, 0.f, 0.f};
  if (m.kr != 0.) {
    f3 wrel = w0 - w1;
    float nwrel = normf(wrel);
    float nwtrel = normf(crossf(w0, -l0) - crossf(w1, -l1));
    if (nwrel > 1e-10f) {
      float Req = 1.f / (1.f / req0 + 1.f / req1);
      dM = wrel * (-((float)m.kr * Req * normFN * nwtrel / nwrel));
    }
  }
  f = delFN + delFT;
  return fn < ft;
}

// ---------------------------------------------------------------------------
// Integrators.  Sphere angular acceleration T_bf / (c I)
// (ParticleKinematicsSphere.cpp:47-55).
struct Kin {
  d3 x, v, w, wb;
  quat q;
};

DEM_HD d3 ang_acc(d3 tbf, double cf, double I) {
  return d3{tbf.x / (cf * I), tbf.y / (cf * I), tbf.z / (cf * I)};
}

// SecondOrderLeapFrog SSPRK3 (SecondOrderLeapFrog.cpp:57-66); for a sphere
// the three stages share one acceleration value.
DEM_HD void ssprk3(d3& vrot, d3 tbf, double cf, double I, double dt) {
  d3 K1 = ang_acc(tbf, cf, I);
  K1 *= dt;
  d3 K2 = ang_acc(tbf, cf, I);
  K2 *= dt;
  d3 K3 = ang_acc(tbf, cf, I);
  K3 *= dt;
  vrot += (K1 + K2 + 4. * K3) / 6.;
}

// Rotation update and translation common to all schemes
// (ParticleKinematics.cpp:78-103); returns |translational motion|.
DEM_HD double finish_move(Kin& k, d3 meanW_bf, d3 tm, double dtx, quat* qdt_out = nullptr) {
  k.w = to_space(k.q, k.wb);
  d3 meanW = to_space(k.q, meanW_bf);
  k.x = k.x + tm;
  double nO = norm(meanW);
  quat qdt;
  if (nO > kEps) {
    double c = cos(nO * dtx / 2.);
    double s = sin(nO * dtx / 2.);
    d3 t = (s * 1. / nO) * meanW;
    qdt = quat{t, c};
  } else
    qdt = quat{d3{0., 0., 0.}, 1.};
  k.q = qmul(qdt, k.q);
  if (qdt_out) *qdt_out = qdt;
  return norm(tm);
}

// ParticleKinematics::Move + SecondOrderLeapFrog::Move (SecondOrderLeapFrog.cpp:44-71)
DEM_HD double lf_move(Kin& k, d3 F, d3 T, double mass, double cf, double I, double dtv, double dtx) {
  d3 tbf = to_body(k.q, T);
  k.v += (F / (mass * cf)) * dtv;
  d3 tm = k.v * dtx;
  ssprk3(k.wb, tbf, cf, I, dtv);
  return finish_move(k, k.wb, tm, dtx);
}
// ParticleKinematics::advanceVelocity + SecondOrderLeapFrog::advanceVelocity
// (ParticleKinematics.cpp:111-128, SecondOrderLeapFrog.cpp:74-95)
DEM_HD void lf_kick(Kin& k, d3 F, d3 T, double mass, double cf, double I, double dtv) {
  d3 tbf = to_body(k.q, T);
  k.v += (F / (mass * cf)) * dtv;
  ssprk3(k.wb, tbf, cf, I, dtv);
  k.w = to_space(k.q, k.wb);
}

// The same for a body of a system with non-spherical particles: bodies
// other than spheres use ParticleKinematics3D (principal moments I3), and
// every body's basis is rotated by the step's quaternion increment
// (Particle::Rotate -> Transform::composeLeftByRotation).
DEM_HD double lf_move_body(Kin& k, bool gyro, d3 F, d3 T, double mass, double cf, double I, d3 I3, double dtv,
                           double dtx, double* B) {
  d3 tbf = to_body(k.q, T);
  k.v += (F / (mass * cf)) * dtv;
  d3 tm = k.v * dtx;
  if (gyro) ssprk3_3(k.wb, tbf, cf, I3, dtv);
  else ssprk3(k.wb, tbf, cf, I, dtv);
  quat qdt;
  double depl = finish_move(k, k.wb, tm, dtx, &qdt);
  compose_left(B, qdt);
  return depl;
}
DEM_HD void lf_kick_body(Kin& k, bool gyro, d3 F, d3 T, double mass, double cf, double I, d3 I3, double dtv) {
  d3 tbf = to_body(k.q, T);
  k.v += (F / (mass * cf)) * dtv;
  if (gyro) ssprk3_3(k.wb, tbf, cf, I3, dtv);
  else ssprk3(k.wb, tbf, cf, I, dtv);
  k.w = to_space(k.q, k.wb);
}

struct AB2State {
  d3 v_nm2, dUdt_nm2, wb_nm2, dWdt_nm2;
};

// FirstOrderExplicit / SecondOrderExplicit / SecondOrderAdamsBashforth Move
// (FirstOrderExplicit.cpp:44-61, SecondOrderExplicit.cpp:44-63,
// SecondOrderAdamsBashforth.cpp:51-101)
DEM_HD double other_move(int scheme, bool first, Kin& k, AB2State& ab, d3 F, d3 T, double mass, double cf,
                         double I, double dtv, double dtx) {
  d3 tbf = to_body(k.q, T);
  d3 tm, meanW;
  if (scheme == kFirstOrderExplicit) {
    tm = k.v * dtx;
    k.v += (F / (mass * cf)) * dtv;
    d3 dw = ang_acc(tbf, cf, I);
    meanW = k.wb;
    k.wb += dw * dtv;
  } else if (scheme == kSecondOrderExplicit) {
    d3 dUdt = F / (mass * cf);
    tm = k.v * dtx + 0.5 * dUdt * dtx * dtx;
    k.v += dUdt * dtv;
    d3 dw = ang_acc(tbf, cf, I);
    meanW = k.wb + 0.5 * dw * dtx;
    k.wb += dw * dtv;
  } else if (first) {
    d3 dUdt = F / (mass * cf);
    tm = k.v * dtx;
    ab.v_nm2 = k.v;
    k.v += dUdt * dtv;
    ab.dUdt_nm2 = dUdt;
    d3 dw = ang_acc(tbf, cf, I);
    meanW = k.wb;
    ab.wb_nm2 = k.wb;
    k.wb += dw * dtv;
    ab.dWdt_nm2 = dw;
  } else {
    d3 dUdt = F / (mass * cf);
    tm = 0.5 * (3. * k.v - ab.v_nm2) * dtx;
    ab.v_nm2 = k.v;
    k.v += 0.5 * (3. * dUdt - ab.dUdt_nm2) * dtv;
    ab.dUdt_nm2 = dUdt;
    d3 dw = ang_acc(tbf, cf, I);
    meanW = 0.5 * (3. * k.wb - ab.wb_nm2);
    ab.wb_nm2 = k.wb;
    k.wb += 0.5 * (3. * dw - ab.dWdt_nm2) * dtv;
    ab.dWdt_nm2 = dw;
  }
  return finish_move(k, meanW, tm, dtx);
}

// ---------------------------------------------------------------------------
// Loads / stores of one particle's state (storage index s)
__device__ __forceinline__ d3 ld3(double* const* a, int s) { return d3{a[0][s], a[1][s], a[2][s]}; }
__device__ __forceinline__ void st3(double* const* a, int s, d3 v) {
  a[0][s] = v.x;
  a[1][s] = v.y;
  a[2][s] = v.z;
}

__device__ __forceinline__ Kin load_kin(const Particles& P, int s, int cx, int cv) {
  Kin k;
  k.x = ld3(P.x[cx], s);
  k.v = ld3(P.v[cv], s);
  k.w = ld3(P.w[cv], s);
  k.wb = ld3(P.wb, s);
  k.q = quat{d3{P.q[0][s], P.q[1][s], P.q[2][s]}, P.q[3][s]};
  return k;
}

// Periodic role swap (LinkedCell.cpp:2004-2310): a particle that drifts into
// the ghost layer is replaced by its periodic clone, i.e. moved by +-L, and
// the clone's body-fixed angular velocity is rebuilt from w and q by
// ParticleKinematics::setAngularVelocity (ParticleKinematics.cpp:197-203).
// Then the cell key is computed and the arrival slot reserved.
__device__ __forceinline__ void wrap_and_bin(Kin& k, const Grid& G, const Sorted& S, DevStatus* st, int s, int pid,
                                             int cell_base = 0, const VerletState* V = nullptr) {
  bool wrapped = false;
  double* xs[3] = {&k.x.x, &k.x.y, &k.x.z};
  int c[3];
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    int ri = ref_cell(*xs[a], G, a);
    if (G.periodic[a]) {
      if (ri < 1) {
        *xs[a] = *xs[a] + G.L[a];
        wrapped = true;
        ri = ref_cell(*xs[a], G, a);
      } else if (ri > G.n[a]) {
        *xs[a] = *xs[a] + (-G.L[a]);
        wrapped = true;
        ri = ref_cell(*xs[a], G, a);
      }
    }
    int gi = ri - G.off[a];
    if (gi < 0 || gi >= G.n[a]) {
      raise_error(st, kSimulationError, pid, 0);
      gi = gi < 0 ? 0 : G.n[a] - 1;
    }
    c[a] = gi;
  }
  if (wrapped) k.wb = to_body(k.q, k.w);
  if (V && V->stamp) {
    // a Verlet list is in use: no binning here (stamp < 0: the list is
    // rebuilt next, nothing to check); the list goes stale once the
    // particle is more than half the skin from where it was built (nearest
    // periodic image, so that a role swap is no displacement)
    if (V->stamp < 0) return;
    double d2 = 0.;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      double d = *xs[a] - V->xb[a][s];
      if (G.periodic[a]) {
        if (d > 0.5 * G.L[a]) d -= G.L[a];
        else if (d < -0.5 * G.L[a]) d += G.L[a];
      }
      d2 += d * d;
    }
    if (d2 > V->half2) st->rebuild = V->stamp;
    return;
  }
  int key = cell_base + (c[1] * G.n[2] + c[2]) * G.n[0] + c[0];  // LinkedCell::getCellNumber order
  S.key[s] = key;
  S.slot[s] = atomicAdd(&S.count[key], 1);
}

// Contact statistics of one thread (AppCollision::addToContactsFeatures,
// AppCollision.cpp:159-189): overlaps of contacts, iterations of the
// closest-point evaluations that report some.
struct StatAcc {
  double sum = 0., max = 0., cnt = 0., git = 0., gn = 0.;
  // one evaluation of overlap d with nit iterations, met mult times
  __device__ __forceinline__ void add(double d, int nit, int mult) {
    if (nit) {
      git += (double)nit * mult;
      gn += mult;
    }
    if (d < 0.) {
      const double o = -d;
      sum += o * mult;
      cnt += mult;
      if (o > max) max = o;
    }
  }
};

// How many times the reference evaluates a particle pair: once for every
// shift u (one periodic length or none per periodic axis) for which both
// A + u and B + u + s exist, B seen at its image B + s -- a particle has a
// clone across a periodic face when it lies in the cell next to the opposite
// face, and clones meet masters and each other in the ghost layers
// (LinkedCell.cpp:1765-2398, :1019-1103).
__device__ __forceinline__ int image_evaluations(const Grid& G, const d3& xa, const d3& xb, const int sh[3]) {
  const double a3[3] = {xa.x, xa.y, xa.z}, b3[3] = {xb.x, xb.y, xb.z};
  int m = 1;
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    if (!G.periodic[a]) continue;
    const int ca = min(max(ref_cell(a3[a], G, a) - G.off[a], 0), G.n[a] - 1);
    const int cb = min(max(ref_cell(b3[a], G, a) - G.off[a], 0), G.n[a] - 1);
    // clone at +L from the first cell, at -L from the last one
    auto exists = [&](int c, int u) { return u == 0 || (u > 0 ? c == 0 : c == G.n[a] - 1); };
    int k = 0;
    for (int u = -1; u <= 1; ++u) {
      const int v = u + sh[a];
      if (v < -1 || v > 1) continue;
      k += exists(ca, u) && exists(cb, v);
    }
    m *= k;
  }
  return m;
}

struct Self {
  d3 x, v, w;
  double r, m, c, req;
  int mat, id;
  StatAcc* acc = nullptr;  // contact statistics, when collected
};

__device__ __forceinline__ void record_pair(const StepParams& sp, DevStatus* st, int* pairs, int a, int b,
                                            bool slip) {
  if (!sp.record_pairs) return;
  if (slip) atomicAdd(&st->n_slip, 1);
  int idx = atomicAdd(&st->npairs, 1);
  if (idx < sp.pair_cap) {
    pairs[2 * idx] = a;
    pairs[2 * idx + 1] = b;
  } else {
    st->pair_overflow = 1;
  }
}

// Body of particle s (storage index) at position x, basis buffer cx.
__device__ __forceinline__ ConvexGeo convex_of(const Particles& P, int s, d3 x, int cx) {
  ConvexGeo g;
  const int sh = P.shape[s];
  for (int j = 0; j < 9; ++j) g.B[j] = P.B[cx][j][s];
  g.o = x;
  if (sh < 0) {
    const double r = P.rad[s], c = P.crust[s];
    g.type = kShapeSphere;
    g.nv = 0;
    g.verts = nullptr;
    g.r = r;
    g.e = d3{0., 0., 0.};
    const double f = r < kEps ? 1. : (r - c) / r;
    g.scaling = d3{f, f, f};
    g.bext = d3{r, r, r};
    g.crust = c;
    g.circR = r;
    g.obb = d3{r, r, r};
  } else {
    const ShapeDesc& d = P.shapes[sh];
    g.type = d.type;
    g.nv = d.nv;
    g.verts = P.verts + 3 * (size_t)d.vert_off;
    g.r = d.radius;
    g.e = d3{d.extent[0], d.extent[1], d.extent[2]};
    g.scaling = d3{d.scaling[0], d.scaling[1], d.scaling[2]};
    g.bext = d3{d.bext[0], d.bext[1], d.bext[2]};
    g.crust = d.crust;
    g.circR = d.circR;
    g.obb = d3{d.obb[0], d.obb[1], d.obb[2]};
  }
  g.opt = P.gjk;
  return g;
}

__device__ __forceinline__ ConvexGeo convex_of_obstacle(const BoxObstacle& B) {
  ConvexGeo g;
  g.type = kShapeBox;
  g.nv = 0;
  g.verts = nullptr;
  g.r = 0.;
  g.e = d3{B.extent[0], B.extent[1], B.extent[2]};
  for (int j = 0; j < 9; ++j) g.B[j] = B.B[j];
  g.o = d3{B.centre[0], B.centre[1], B.centre[2]};
  const double* e = B.extent;
  g.scaling = d3{e[0] < kEps ? 1. : (e[0] - B.crust) / e[0], e[1] < kEps ? 1. : (e[1] - B.crust) / e[1],
                 e[2] < kEps ? 1. : (e[2] - B.crust) / e[2]};
  g.bext = g.e;
  g.crust = B.crust;
  g.circR = norm(g.e);
  g.obb = g.e;
  g.opt = GjkOpts{0, 0, 0, 0, 1e-10};  // the options are the particle's (body A of every pair)
  return g;
}

// Periodic image offset of a neighbour seen across a face: -L, 0 or +L on
// axis a for image code -1, 0, +1.
__device__ __forceinline__ double image_shift(const Grid& G, int a, int c) {
  return c < 0 ? -G.L[a] : (c > 0 ? G.L[a] : 0.);
}

// Enumerates the neighbour candidates of a particle at x as runs of sorted
// slots [b, e) of its 27 neighbour cells, each with the periodic image
// (ix, iy, iz) under which its particles are seen, in a fixed order (rows
// (dj, dk) of x-adjacent cells, each row's unshifted run, then its cells
// across the periodic x faces).  Cells of one row are consecutive keys, so
// each row is one contiguous run of sorted slots; the nine row ranges are
// loaded before any candidate.
template <class F>
__device__ __forceinline__ void for_each_run(const Sorted& S, const Grid& G, int start_base, int slot_lo,
                                             int slot_hi, d3 x, F&& f) {
  // own cell, clamped so that a particle flagged outside the grid cannot
  // index outside the cell table
  const int ci = min(max(ref_cell(x.x, G, 0) - G.off[0], 0), G.n[0] - 1);
  const int cj = min(max(ref_cell(x.y, G, 1) - G.off[1], 0), G.n[1] - 1);
  const int ck = min(max(ref_cell(x.z, G, 2) - G.off[2], 0), G.n[2] - 1);
  int i0 = ci - 1, i1 = ci + 1;
  bool wrap_lo = false, wrap_hi = false;
  if (G.periodic[0]) {
    if (i0 < 0) { wrap_lo = true; i0 = 0; }
    if (i1 >= G.n[0]) { wrap_hi = true; i1 = G.n[0] - 1; }
  } else {
    i0 = max(i0, 0);
    i1 = min(i1, G.n[0] - 1);
  }
  // neighbour row r = 3 (dj + 1) + (dk + 1): wrapped cell indices, image
  // codes on y and z, and whether the row exists
  auto row_of = [&](int r, int& row, int& iy, int& iz) {
    int nj = cj + r / 3 - 1, nk = ck + r % 3 - 1;
    iy = 0;
    iz = 0;
    if (G.periodic[1]) {
      if (nj < 0) { nj += G.n[1]; iy = -1; }
      else if (nj >= G.n[1]) { nj -= G.n[1]; iy = 1; }
    } else if (nj < 0 || nj >= G.n[1]) return false;
    if (G.periodic[2]) {
      if (nk < 0) { nk += G.n[2]; iz = -1; }
      else if (nk >= G.n[2]) { nk -= G.n[2]; iz = 1; }
    } else if (nk < 0 || nk >= G.n[2]) return false;
    row = start_base + (nj * G.n[2] + nk) * G.n[0];
    return true;
  };
  int rb[9], re[9];
#pragma unroll
  for (int r = 0; r < 9; ++r) {
    int row, iy, iz;
    const bool ok = row_of(r, row, iy, iz);
    rb[r] = ok ? max(S.start[row + i0], slot_lo) : 0;
    re[r] = ok ? min(S.start[row + i1 + 1], slot_hi) : 0;
  }
  for (int r = 0; r < 9; ++r) {
    int row, iy, iz;
    if (!row_of(r, row, iy, iz)) continue;
    // the row's unshifted run, then its cells across the periodic x faces
    // (one call site of f, so that the visitor is inlined once)
    for (int seg = 0; seg < 3; ++seg) {
      int b, e, ix;
      if (seg == 0) {
        b = rb[r];
        e = re[r];
        ix = 0;
      } else {
        if (seg == 1 ? !wrap_lo : !wrap_hi) continue;
        const int c0 = seg == 1 ? row + G.n[0] - 1 : row;
        b = max(S.start[c0], slot_lo);
        e = min(S.start[c0 + 1], slot_hi);
        ix = seg == 1 ? -1 : 1;
      }
      if (b < e) f(b, e, ix, iy, iz);
    }
  }
}

// The candidates of for_each_run one by one: f(k2, ix, iy, iz).
template <class F>
__device__ __forceinline__ void for_each_candidate(const Sorted& S, const Grid& G, int start_base, int slot_lo,
                                                   int slot_hi, d3 x, F&& f) {
  for_each_run(S, G, start_base, slot_lo, slot_hi, x, [&](int b, int e, int ix, int iy, int iz) {
    for (int k2 = b; k2 < e; ++k2) f(k2, ix, iy, iz);
  });
}

// Narrow phase of a neighbour candidate that passed the conservative
// distance reject: particle s2i (storage index; sorted slot k2 when known,
// id pid2, radius ro) seen at xo, AB = xo - x_self, nab = Norm(AB); for a
// contact (overlap < 0), the visitor.  Particle::InterAction
// (Particle.cpp:477-545) with the ContactError recovery of
// RigidBodyWithCrust.cpp:534-590.
// In the statistics a pair counts on the side of its lower id, as often as
// the reference meets it (image_evaluations; sh = the image code of xo).
template <class V, bool SHAPES>
__device__ __forceinline__ void narrow_visit(const Particles& P, const Grid& G, DevStatus* st, const Self& me,
                                             V& vis, const ConvexGeo* mg, int cx, int k2, int s2i, int pid2, d3 xo,
                                             double ro, d3 AB, double nab, const d3& xb, const int sh[3]) {
  const double co = P.crust[s2i];
  d3 Pc, ov;
  double d;
  int nit = 1;
  if (SHAPES && (mg->type != kShapeSphere ||
                 (P.shape[s2i] >= 0 && P.shapes[P.shape[s2i]].type != kShapeSphere))) {
    ConvexGeo og = convex_of(P, s2i, xo, cx);
    int res = closest_convex(*mg, og, Pc, ov, d, &nit);
    if (res == 0) return;
    if (res < 0) {
      if (!recover_gjk(*mg, og, Pc, ov, d, &nit)) {
        raise_error(st, kContactError, me.id, pid2);
        return;
      }
      atomicAdd(&st->n_recovered, 1);
    }
  } else {
    int res = sphere_sphere(me.x, me.r, me.c, AB, nab, ro, co, Pc, ov, d);
    if (res == 0) return;
    if (res < 0) {
      if (!recover_ss(me.x, me.r, me.c, xo, ro, co, Pc, ov, d)) {
        raise_error(st, kContactError, me.id, pid2);
        return;
      }
      atomicAdd(&st->n_recovered, 1);
    }
  }
  if (me.acc && me.id < pid2) me.acc->add(d, nit, image_evaluations(G, me.x, xb, sh));
  if (!(d < 0.)) return;
  vis.pair(k2, s2i, pid2, xo, ro, Pc, ov, d);
}

// One neighbour candidate of the cell sweep (sorted slot k2, periodic image
// (ix, iy, iz)).
template <class V, bool SHAPES>
__device__ __forceinline__ void visit_candidate(const Particles& P, const Sorted& S, const Grid& G, DevStatus* st,
                                                int slot_lo, int slot_hi, const Self& me, V& vis,
                                                const ConvexGeo* mg, int cx, int k2, int ix, int iy, int iz) {
  const double* q2 = S.posr + 4 * (size_t)k2;
  // periodic clone position x + vec (added only on the wrapped axes)
  d3 xo{ix ? q2[0] + image_shift(G, 0, ix) : q2[0], iy ? q2[1] + image_shift(G, 1, iy) : q2[1],
        iz ? q2[2] + image_shift(G, 2, iz) : q2[2]};
  double ro = q2[3];
  d3 AB = xo - me.x;
  double s2 = AB.x * AB.x + AB.y * AB.y + AB.z * AB.z;
  double R = me.r + ro;
  // conservative reject: s2 > R^2 (1 + 1e-12) implies sqrt(s2) - R > 0
  if (s2 > R * R * 1.000000000001) return;
  double nab = sqrt(s2);
  const int s2i = S.order[k2];
  if (s2i < slot_lo || s2i >= slot_hi) return;
  const int sh[3] = {ix, iy, iz};
  narrow_visit<V, SHAPES>(P, G, st, me, vis, mg, cx, k2, s2i, S.sid[k2], xo, ro, AB, nab, d3{q2[0], q2[1], q2[2]},
                          sh);
}

// Particle s2i at its current position (position buffer cx), seen at its
// periodic image nearest to this particle: for any pair closer than half
// the domain -- every pair in contact -- that is the clone position x + vec
// of the cell sweep, bit for bit.
__device__ __forceinline__ d3 nearest_image(const Grid& G, d3 xo, const d3& x) {
  double* xs[3] = {&xo.x, &xo.y, &xo.z};
  const double* ms[3] = {&x.x, &x.y, &x.z};
#pragma unroll
  for (int a = 0; a < 3; ++a)
    if (G.periodic[a]) {
      const double d = *xs[a] - *ms[a];
      if (d > 0.5 * G.L[a]) *xs[a] = *xs[a] + image_shift(G, a, -1);
      else if (d < -0.5 * G.L[a]) *xs[a] = *xs[a] + image_shift(G, a, 1);
    }
  return xo;
}

template <class V, bool SHAPES>
__device__ __forceinline__ void visit_current_at(const Particles& P, const Grid& G, DevStatus* st, const Self& me,
                                                 V& vis, const ConvexGeo* mg, int cx, int s2i, d3 x2, double ro) {
  const d3 xo = nearest_image(G, x2, me.x);
  d3 AB = xo - me.x;
  double s2 = AB.x * AB.x + AB.y * AB.y + AB.z * AB.z;
  double R = me.r + ro;
  if (s2 > R * R * 1.000000000001) return;
  const int sh[3] = {xo.x < x2.x ? -1 : (xo.x > x2.x ? 1 : 0), xo.y < x2.y ? -1 : (xo.y > x2.y ? 1 : 0),
                     xo.z < x2.z ? -1 : (xo.z > x2.z ? 1 : 0)};
  narrow_visit<V, SHAPES>(P, G, st, me, vis, mg, cx, -1, s2i, P.id[s2i], xo, ro, AB, sqrt(s2), x2, sh);
}

template <class V, bool SHAPES>
__device__ __forceinline__ void visit_current(const Particles& P, const Grid& G, DevStatus* st, const Self& me,
                                              V& vis, const ConvexGeo* mg, int cx, int s2i) {
  visit_current_at<V, SHAPES>(P, G, st, me, vis, mg, cx, s2i, d3{P.x[cx][0][s2i], P.x[cx][1][s2i], P.x[cx][2][s2i]},
                              P.rad[s2i]);
}

// Particle-obstacle contacts (SimpleObstacle::InterAction,
// SimpleObstacle.cpp:222-271).
template <class V, bool SHAPES>
__device__ __forceinline__ void visit_obstacles(const BoxObstacle* __restrict__ obs, int nobs, DevStatus* st,
                                                const Self& me, V& vis, const ConvexGeo* mg) {
  for (int o = 0; o < nobs; ++o) {
    const BoxObstacle& B = obs[o];
    bool probable = fabs(B.aabb_c[0] - me.x.x) <= B.aabb_e[0] + me.r &&
                    fabs(B.aabb_c[1] - me.x.y) <= B.aabb_e[1] + me.r &&
                    fabs(B.aabb_c[2] - me.x.z) <= B.aabb_e[2] + me.r;
    if (!probable) continue;
    d3 Pc, ov;
    double d;
    int nit = 1;
    if (SHAPES && mg->type != kShapeSphere) {
      ConvexGeo og = convex_of_obstacle(B);
      int res = closest_convex(*mg, og, Pc, ov, d, &nit);
      if (res == 0) continue;
      if (res < 0) {
        if (!recover_gjk(*mg, og, Pc, ov, d, &nit)) {
          raise_error(st, kContactError, me.id, B.id);
          continue;
        }
        atomicAdd(&st->n_recovered, 1);
      }
    } else {
      int res = sphere_box(B, me.x, me.r, me.c, Pc, ov, d);
      if (res == 0) continue;
      if (res < 0) {
        if (!recover_sb(B, me.x, me.r, me.c, Pc, ov, d)) {
          raise_error(st, kContactError, me.id, B.id);
          continue;
        }
        atomicAdd(&st->n_recovered, 1);
      }
    }
    if (me.acc) me.acc->add(d, nit, 1);
    if (!(d < 0.)) continue;
    vis.wall(o, Pc, ov, d);
  }
}

// Enumerates the contacts (overlap < 0) of one particle (sorted slot k) with
// its 27 neighbour cells and the obstacles -- LinkedCell::ComputeForces +
// Particle::InterAction + SimpleObstacle::InterAction (LinkedCell.cpp:1019-1103,
// Particle.cpp:477-545, SimpleObstacle.cpp:222-271) -- and hands each one to
// the visitor:
//   vis.pair(k2, s2i, xo, Pc, ov, d)   neighbour in sorted slot k2, storage
//                                      s2i, seen at position xo (periodic image)
//   vis.wall(o, Pc, ov, d)             obstacle o
// with the contact point Pc, overlap vector ov and overlap distance d < 0 of
// RigidBodyWithCrust::ClosestPoint (body A = this particle).  start_base
// offsets the cell-start table; [slot_lo, slot_hi) bounds the sorted slots and
// storage indices of the particle's system (one system: 0 and [0, N)).
// With SHAPES the system holds non-spherical particles: mg describes this
// particle's body (its shape, basis and crust) and cx selects the basis
// buffer of the neighbours.
template <class V, bool SHAPES = false>
__device__ __forceinline__ void for_each_contact(const Particles& P, const Sorted& S, const Grid& G,
                                                 const BoxObstacle* __restrict__ obs, int nobs, DevStatus* st,
                                                 int start_base, int slot_lo, int slot_hi, int k, const Self& me,
                                                 V& vis, const ConvexGeo* mg = nullptr, int cx = 0) {
  for_each_candidate(S, G, start_base, slot_lo, slot_hi, me.x, [&](int k2, int ix, int iy, int iz) {
    if (k2 != k) visit_candidate<V, SHAPES>(P, S, G, st, slot_lo, slot_hi, me, vis, mg, cx, k2, ix, iy, iz);
  });
  visit_obstacles<V, SHAPES>(obs, nobs, st, me, vis, mg);
}

// Candidate entry of a neighbour list: sorted slot in the low 27 bits,
// periodic image code (ix+1) + 3(iy+1) + 9(iz+1) in the high 5.
constexpr int kNbrSlotBits = kNbrSlotBitsHost;
constexpr unsigned kNbrSlotMask = (1u << kNbrSlotBits) - 1u;
__device__ __forceinline__ unsigned nbr_entry(int k2, int ix, int iy, int iz) {
  return (unsigned)k2 | ((unsigned)((ix + 1) + 3 * (iy + 1) + 9 * (iz + 1)) << kNbrSlotBits);
}
constexpr int kNbrNoShift = 13;  // the image code of a neighbour without periodic shift

// The same contacts as for_each_contact, from the candidate list written by
// k_detect for this step (same candidates, same order).  A particle whose
// candidates overflowed the list is swept in full.
template <class V, bool SHAPES = false>
__device__ __forceinline__ void for_each_listed(const Particles& P, const Sorted& S, const Grid& G,
                                                const BoxObstacle* __restrict__ obs, int nobs, DevStatus* st,
                                                int slot_lo, int slot_hi, int k, const Self& me, V& vis,
                                                const NbrList& L, const ConvexGeo* mg = nullptr, int cx = 0,
                                                int start_base = 0) {
  const int c = L.count[k];
  if (c > L.cap) {
    for_each_contact<V, SHAPES>(P, S, G, obs, nobs, st, start_base, slot_lo, slot_hi, k, me, vis, mg, cx);
    return;
  }
  for (int i = 0; i < c; ++i) {
    const unsigned e = L.ent[(size_t)i * L.n + k];
    const int code = (int)(e >> kNbrSlotBits);
    visit_candidate<V, SHAPES>(P, S, G, st, slot_lo, slot_hi, me, vis, mg, cx, (int)(e & kNbrSlotMask),
                               code % 3 - 1, (code / 3) % 3 - 1, code / 9 - 1);
  }
  visit_obstacles<V, SHAPES>(obs, nobs, st, me, vis, mg);
}

// The contacts of particle s from the Verlet list (neighbours by storage
// index, in the order of the cell sweep at the last build); a particle
// whose candidates overflowed the list enumerates the same cells of the
// list's grid around its build position, in the same order.
template <class V, bool SHAPES = false>
__device__ __forceinline__ void for_each_verlet(const Particles& P, const Sorted& S, const Grid& G,
                                                const VerletState& VS, const BoxObstacle* __restrict__ obs, int nobs,
                                                DevStatus* st, int n, int s, const Self& me, V& vis,
                                                const NbrList& L, const ConvexGeo* mg = nullptr, int cx = 0,
                                                const NbrList* CL = nullptr) {
  // this step's contacts, filtered from the list (k_vfilter) in its order
  const int c2 = CL ? CL->count[s] : 0;
  if (CL && c2 <= CL->cap) {
    for (int i = 0; i < c2; ++i) {
      const unsigned s2i = CL->ent[(size_t)i * CL->n + s];
      if (s2i < (unsigned)n) visit_current<V, SHAPES>(P, G, st, me, vis, mg, cx, (int)s2i);
    }
    visit_obstacles<V, SHAPES>(obs, nobs, st, me, vis, mg);
    return;
  }
  const int c = L.count[s];
  if (c <= L.cap) {
    for (int i = 0; i < c; ++i) {
      const unsigned s2i = L.ent[(size_t)i * L.n + s];
      if (s2i < (unsigned)n) visit_current<V, SHAPES>(P, G, st, me, vis, mg, cx, (int)s2i);
    }
  } else {
    const d3 xb{VS.xb[0][s], VS.xb[1][s], VS.xb[2][s]};
    for_each_candidate(S, VS.Gv, 0, 0, n, xb, [&](int k2, int, int, int) {
      const int s2i = S.order[k2];
      if (s2i != s && (unsigned)s2i < (unsigned)n) visit_current<V, SHAPES>(P, G, st, me, vis, mg, cx, s2i);
    });
  }
  visit_obstacles<V, SHAPES>(obs, nobs, st, me, vis, mg);
}

// Hooke-dashpot-Coulomb force and torque gathered on the particle
// (HookeContactForceModel.cpp:121-163, Torsor.cpp:184-190); F enters holding
// the weight.  MIXED evaluates the model in FP32 on lever arms.
// EXTRAS: contact memory models and statistics may occur (checked at run
// time); without, the visitor carries none of their code.
template <bool MIXED, bool EXTRAS = false>
struct ForceVisitor {
  const Particles& P;
  const Sorted& S;
  const BoxObstacle* obs;
  const HookeModel* models;
  const StepParams& sp;
  DevStatus* st;
  int* pairs;
  int cv;
  const Self& me;
  d3 F, T;
  f3 Fc, Tc;
  int s = -1;     // storage index of self (contact memory)
  int nmem = 0;   // memory entries written by this evaluation
  double fs = 0.; // sum of the contact-force magnitudes (EXTRAS with sp.fsum)

  // HookeMemory: this pair's entry of the last evaluation (zero when new),
  // the force, and this evaluation's entry.  Evaluated in FP64 in both
  // precisions.
  __device__ __forceinline__ bool memory_pair(const HookeModel& hm, int pid2, d3 xo, d3 vo, d3 wo, double mo,
                                              double reqo, d3 Pc, d3 ov) {
    const ContactMemory& cm = sp.mem;
    const size_t n = (size_t)sp.n;
    const int c = cm.cur, o = c ^ 1;
    int j = -1;
    for (int i = 0, cnt = cm.n[c][s]; i < cnt; ++i)
      if (cm.key[c][i * n + s] == pid2) {
        j = i;
        break;
      }
    d3 kd{0., 0., 0.}, pn{0., 0., 0.}, sr{0., 0., 0.};
    if (j >= 0) {
      const size_t e = j * n + s;
      kd = d3{cm.kd[c][0][e], cm.kd[c][1][e], cm.kd[c][2][e]};
      pn = d3{cm.pn[c][0][e], cm.pn[c][1][e], cm.pn[c][2][e]};
      sr = d3{cm.sr[c][0][e], cm.sr[c][1][e], cm.sr[c][2][e]};
    }
    d3 f, dM;
    const bool slip = hooke_memory64(hm, me.x, me.v, me.w, me.m, me.req, xo, vo, wo, mo, reqo, Pc, ov, sp.dt, j >= 0,
                                     kd, pn, sr, f, dM);
    const int jn = nmem++;
    if (jn < cm.cap) {
      const size_t e = jn * n + s;
      cm.key[o][e] = pid2;
      cm.kd[o][0][e] = kd.x;
      cm.kd[o][1][e] = kd.y;
      cm.kd[o][2][e] = kd.z;
      cm.pn[o][0][e] = pn.x;
      cm.pn[o][1][e] = pn.y;
      cm.pn[o][2][e] = pn.z;
      cm.sr[o][0][e] = sr.x;
      cm.sr[o][1][e] = sr.y;
      cm.sr[o][2][e] = sr.z;
    }
    F += f;
    T += cross(Pc - me.x, f);
    if (hm.mur != 0.) T += dM * 1.;
    if (sp.fsum) fs += norm(f);
    return slip;
  }

  __device__ __forceinline__ void pair(int k2, int s2i, int pid2, d3 xo, double ro, d3 Pc, d3 ov, double d) {
    (void)k2;
    (void)ro;
    (void)d;
    const HookeModel& hm = models[me.mat * kMaxMaterials + P.mat[s2i]];
    d3 vo = ld3(P.v[cv], s2i), wo = ld3(P.w[cv], s2i);
    double mo = P.mass[s2i], reqo = P.req[s2i];
    bool slip;
    if (EXTRAS && hm.memory) {
      slip = memory_pair(hm, pid2, xo, vo, wo, mo, reqo, Pc, ov);
    } else if (MIXED) {
      f3 l0 = f3of(Pc - me.x), l1 = f3of(Pc - xo);
      f3 f, dM;
      slip = hooke32(hm, l0, f3of(me.w), (float)me.m, (float)me.req, l1, f3of(wo), (float)mo, (float)reqo,
                     f3of(rel_velocity(me.x, me.v, me.w, xo, vo, wo, Pc)), f3of(ov), f, dM);
      Fc += f;
      Tc += crossf(l0, f);
      if (hm.kr != 0.) Tc += dM;
      if (EXTRAS && sp.fsum) fs += (double)normf(f);
    } else {
      d3 f, dM;
      slip = hooke64(hm, me.x, me.v, me.w, me.m, me.req, xo, vo, wo, mo, reqo, Pc, ov, f, dM);
      F += f;
      T += cross(Pc - me.x, f);
      if (hm.kr != 0.) T += dM * 1.;
      if (EXTRAS && sp.fsum) fs += norm(f);
    }
    record_pair(sp, st, pairs, me.id, pid2, slip);
  }

  __device__ __forceinline__ void wall(int o, d3 Pc, d3 ov, double d) {
    (void)d;
    const BoxObstacle& B = obs[o];
    const HookeModel& hm = models[me.mat * kMaxMaterials + B.mat];
    d3 cen{B.centre[0], B.centre[1], B.centre[2]};
    d3 zero{0., 0., 0.};
    bool slip;
    if (EXTRAS && hm.memory) {
      slip = memory_pair(hm, B.id, cen, zero, zero, kObstacleMass, B.req, Pc, ov);
    } else if (MIXED) {
      f3 l0 = f3of(Pc - me.x), l1 = f3of(Pc - cen);
      f3 f, dM;
      slip = hooke32(hm, l0, f3of(me.w), (float)me.m, (float)me.req, l1, f3{0.f, 0.f, 0.f}, (float)kObstacleMass,
                     (float)B.req, f3of(rel_velocity(me.x, me.v, me.w, cen, zero, zero, Pc)), f3of(ov), f, dM);
      Fc += f;
      Tc += crossf(l0, f);
      if (hm.kr != 0.) Tc += dM;
      if (EXTRAS && sp.fsum) fs += (double)normf(f);
    } else {
      d3 f, dM;
      slip = hooke64(hm, me.x, me.v, me.w, me.m, me.req, cen, zero, zero, kObstacleMass, B.req, Pc, ov, f, dM);
      F += f;
      T += cross(Pc - me.x, f);
      if (hm.kr != 0.) T += dM * 1.;
      if (EXTRAS && sp.fsum) fs += norm(f);
    }
    record_pair(sp, st, pairs, me.id, B.id, slip);
  }
};

template <bool MIXED, bool SHAPES = false, bool EXTRAS = false>
__device__ __forceinline__ void sweep_particle(const Particles& P, const Sorted& S, const Grid& G,
                                               const BoxObstacle* __restrict__ obs, int nobs,
                                               const HookeModel* __restrict__ models, const StepParams& sp,
                                               DevStatus* st, int* pairs, int start_base, int slot_lo, int slot_hi,
                                               int cv, int k, const Self& me, d3& F, d3& T,
                                               const ConvexGeo* mg = nullptr, int cx = 0,
                                               const NbrList* L = nullptr, const VerletState* VS = nullptr,
                                               int s = -1, const NbrList* CL = nullptr) {
  using FV = ForceVisitor<MIXED, EXTRAS>;
  FV vis{P, S, obs, models, sp, st, pairs, cv, me, F, T, f3{0.f, 0.f, 0.f}, f3{0.f, 0.f, 0.f}};
  vis.s = s;
  if (VS && VS->stamp)
    for_each_verlet<FV, SHAPES>(P, S, G, *VS, obs, nobs, st, slot_hi, s, me, vis, *L, mg, cx,
                                CL && CL->count ? CL : nullptr);
  else if (L && L->count)
    for_each_listed<FV, SHAPES>(P, S, G, obs, nobs, st, slot_lo, slot_hi, k, me, vis, *L, mg, cx, start_base);
  else
    for_each_contact<FV, SHAPES>(P, S, G, obs, nobs, st, start_base, slot_lo, slot_hi, k, me, vis, mg, cx);
  F = vis.F;
  T = vis.T;
  if (MIXED) {
    F += d3{(double)vis.Fc.x, (double)vis.Fc.y, (double)vis.Fc.z};
    T += d3{(double)vis.Tc.x, (double)vis.Tc.y, (double)vis.Tc.z};
  }
  if (EXTRAS && sp.fsum && s >= 0) sp.fsum[s] = vis.fs;
  // this evaluation's contact memory of the particle (every particle, so
  // that entries of contacts that ended are dropped)
  if (EXTRAS && sp.mem.cap && s >= 0) {
    sp.mem.n[sp.mem.cur ^ 1][s] = min(vis.nmem, sp.mem.cap);
    if (vis.nmem > sp.mem.cap) raise_error(st, kMemoryOverflow, me.id, vis.nmem);
  }
}

// Orders the particles of one cell (sorted slots [b, e)) by id, writes
// order / sid / posr for them.  Storage indices outside [lo, hi) (only after
// an interrupted binning) are replaced by lo so that no access leaves the
// arrays.
constexpr int kCellLocal = 16;
__device__ __forceinline__ void canon_cell(const Particles& P, const Sorted& S, int b, int e, int cx, int lo,
                                           int hi) {
  int cnt = e - b;
  if (cnt <= 0) return;
  if (cnt <= kCellLocal) {
    int ent[kCellLocal], ids[kCellLocal];
    for (int i = 0; i < cnt; ++i) {
      int s = S.tmp_order[b + i];
      if (s < lo || s >= hi) s = lo;
      int id = P.id[s];
      int j = i;
      while (j > 0 && ids[j - 1] > id) {
        ids[j] = ids[j - 1];
        ent[j] = ent[j - 1];
        --j;
      }
      ids[j] = id;
      ent[j] = s;
    }
    for (int i = 0; i < cnt; ++i) {
      int s = ent[i];
      S.order[b + i] = s;
      S.sid[b + i] = ids[i];
      if (S.inv) S.inv[s] = b + i;
      double* pr = S.posr + 4 * (size_t)(b + i);
      pr[0] = P.x[cx][0][s];
      pr[1] = P.x[cx][1][s];
      pr[2] = P.x[cx][2][s];
      pr[3] = P.rad[s];
    }
  } else {
    // crowded cell: insertion sort in global memory
    for (int i = 0; i < cnt; ++i) {
      int s = S.tmp_order[b + i];
      if (s < lo || s >= hi) s = lo;
      int id = P.id[s];
      int j = i;
      while (j > 0 && S.sid[b + j - 1] > id) {
        S.sid[b + j] = S.sid[b + j - 1];
        S.order[b + j] = S.order[b + j - 1];
        --j;
      }
      S.sid[b + j] = id;
      S.order[b + j] = s;
    }
    for (int i = 0; i < cnt; ++i) {
      int s = S.order[b + i];
      if (S.inv) S.inv[s] = b + i;
      double* pr = S.posr + 4 * (size_t)(b + i);
      pr[0] = P.x[cx][0][s];
      pr[1] = P.x[cx][1][s];
      pr[2] = P.x[cx][2][s];
      pr[3] = P.rad[s];
    }
  }
}

}  // namespace dem
