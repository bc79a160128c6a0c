// Reverse-mode derivative (vector-Jacobian product) of one Hooke-dashpot-
// Coulomb contact, written by hand.  The forward expressions are those of
// HookeContactForceModel::performForcesCalculus (HookeContactForceModel.cpp:44-114)
// with the closed-form sphere-sphere / sphere-box geometry of
// RigidBodyWithCrust::ClosestPoint (RigidBodyWithCrust.cpp:596-720, Box.cpp:466-769)
// restated in exact arithmetic:
//   u  = unit vector from this sphere's centre towards the other body
//   d  = overlap distance (< 0),   ov = d u,   normal n = -u (|n_k| < 1e-10 -> 0)
//   l0 = P - x0 = (r0 + kappa d) u          (kappa = 1/2; 3/2 for box edges and
//                                           corners, whose contact point lies on
//                                           the sphere side, Box.cpp:725-769)
//   l1 = P - x1 = -(r1 + d/2) u             (other sphere)
//   tv = (v0 + w0 x l0) - (v1 + w1 x l1),  vn = (tv.n) n,  vt = tv - vn
//   FN = kn ov - gn vn,  gn = -2 beta sqrt(m* kn)
//   FT = -gt vt,  gt = 2 et m*,  et = -beta sqrt(kn / m*)   (or a given etat)
//   Coulomb: if mu |FN| < |FT|:  FT = -mu |FN| vt / |vt|   (0 if |vt| <= 1e-10)
//   f = FN + FT on body 0, -f on body 1; torques l0 x f and l1 x (-f)
// For the cotangents (F0b, T0b) of body 0's force/torque and (F1b, T1b) of
// body 1's, the scalar L = F0b.f + T0b.(l0 x f) - F1b.f - T1b.(l1 x f) is
// differentiated with respect to body 0's position, velocity and angular
// velocity and to (kn, beta, mu).  Rolling resistance (kr) is not
// differentiated.  The branch decisions (contact, Coulomb switch, |vt|
// threshold, normal rounding) are those of the forward evaluation.
#pragma once
#include "dem_math.hpp"

namespace dem {

struct PairGrad {
  d3 x0, v0, w0;           // dL / d(body-0 state)
  double kn, beta, mu;     // dL / d(parameters)
};

struct ContactIn {
  d3 u;          // unit vector body 0 -> body 1 (feature)
  double d;      // overlap distance (< 0)
  double r0;     // radius of body 0
  double kappa;  // contact-point factor of body 0's lever
  bool sphere1;  // body 1 is a sphere (else a static obstacle)
  double r1;     // radius of body 1 (spheres)
  d3 v0, w0, v1, w1;
  double m0, m1;
  double kn, beta, etat, mu;
};

// Forward values needed by the backward pass
struct ContactFwd {
  d3 l0, l1, n, tv, vn, vt, FN, FT0, FT, f;
  double vns, nFN, mstar, sq1, sq2, et, gn, gt, nvt, c;
  bool slip, vt_small;
  int nmask[3];
};

DEM_HD void contact_forward(const ContactIn& in, ContactFwd& w) {
  const d3 u = in.u;
  w.l0 = (in.r0 + in.kappa * in.d) * u;
  w.l1 = in.sphere1 ? (-(in.r1 + 0.5 * in.d)) * u : d3{0., 0., 0.};
  d3 n = -u;
  w.nmask[0] = fabs(n.x) < kEps ? 0 : 1;
  w.nmask[1] = fabs(n.y) < kEps ? 0 : 1;
  w.nmask[2] = fabs(n.z) < kEps ? 0 : 1;
  w.n = d3{w.nmask[0] ? n.x : 0., w.nmask[1] ? n.y : 0., w.nmask[2] ? n.z : 0.};
  w.tv = (in.v0 + cross(in.w0, w.l0)) - (in.v1 + cross(in.w1, w.l1));
  w.vns = dot(w.tv, w.n);
  w.vn = w.vns * w.n;
  w.vt = w.tv - w.vn;
  w.mstar = 1. / (1. / in.m0 + 1. / in.m1);
  w.sq1 = sqrt(w.mstar * in.kn);
  w.gn = -2. * in.beta * w.sq1;
  d3 ov = in.d * u;
  w.FN = in.kn * ov - w.gn * w.vn;
  w.nFN = norm(w.FN);
  if (in.etat == -1.) {
    w.sq2 = sqrt(in.kn / w.mstar);
    w.et = -in.beta * w.sq2;
  } else {
    w.sq2 = 0.;
    w.et = in.etat;
  }
  w.gt = 2. * w.et * w.mstar;
  w.FT0 = -w.gt * w.vt;
  double fn = in.mu * w.nFN, ft = norm(w.FT0);
  w.slip = fn < ft;
  w.nvt = norm(w.vt);
  w.vt_small = !(w.nvt > kEps);
  w.c = 0.;
  if (w.slip) {
    if (!w.vt_small) {
      w.c = -in.mu * w.nFN / w.nvt;
      w.FT = w.c * w.vt;
    } else {
      w.FT = d3{0., 0., 0.};
    }
  } else {
    w.FT = w.FT0;
  }
  w.f = w.FN + w.FT;
}

// Backward pass.  Returns dL/du and dL/dd (geometry cotangents) besides the
// state and parameter gradients; the caller maps (u, d) back to positions.
DEM_HD void contact_backward(const ContactIn& in, const ContactFwd& w, d3 F0b, d3 T0b, d3 F1b, d3 T1b,
                             PairGrad& g, d3& g_u, double& g_d) {
  const d3 u = in.u;
  d3 a = F0b - F1b;
  d3 g_f = a + cross(T0b, w.l0);
  d3 g_l0 = cross(w.f, T0b);
  d3 g_l1{0., 0., 0.};
  if (in.sphere1) {
    g_f = g_f - cross(T1b, w.l1);
    g_l1 = -cross(w.f, T1b);
  }
  d3 g_FN = g_f, g_FT = g_f;
  d3 g_vt{0., 0., 0.};
  double g_nFN = 0., g_gt = 0.;
  g.mu = 0.;
  if (w.slip) {
    if (!w.vt_small) {
      g_vt = w.c * g_FT;
      double g_c = dot(g_FT, w.vt);
      g.mu += g_c * (-w.nFN / w.nvt);
      g_nFN += g_c * (-in.mu / w.nvt);
      double g_nvt = g_c * (in.mu * w.nFN / (w.nvt * w.nvt));
      g_vt += (g_nvt / w.nvt) * w.vt;
    }
  } else {
    g_vt = (-w.gt) * g_FT;
    g_gt = -dot(g_FT, w.vt);
  }
  if (w.nFN > 0.) g_FN += (g_nFN / w.nFN) * w.FN;
  // FN = kn d u - gn vn
  g.kn = dot(g_FN, in.d * u);
  d3 g_ov = in.kn * g_FN;
  double g_gn = -dot(g_FN, w.vn);
  d3 g_vn = (-w.gn) * g_FN;
  // gt = 2 et m*;  et = -beta sqrt(kn/m*)
  double g_et = 2. * w.mstar * g_gt;
  g.beta = 0.;
  if (in.etat == -1.) {
    g.beta += -w.sq2 * g_et;
    double g_sq2 = -in.beta * g_et;
    g.kn += g_sq2 * 0.5 / (w.sq2 * w.mstar);
  }
  // gn = -2 beta sqrt(m* kn)
  g.beta += -2. * w.sq1 * g_gn;
  double g_sq1 = -2. * in.beta * g_gn;
  g.kn += g_sq1 * 0.5 * w.mstar / w.sq1;
  // vt = tv - vn;  vn = vns n;  vns = tv.n
  d3 g_tv = g_vt;
  g_vn -= g_vt;
  double g_vns = dot(g_vn, w.n);
  d3 g_n = w.vns * g_vn;
  g_tv += g_vns * w.n;
  g_n += g_vns * w.tv;
  // tv = (v0 + w0 x l0) - (v1 + w1 x l1)
  g.v0 = g_tv;
  g.w0 = cross(w.l0, g_tv);
  g_l0 += cross(g_tv, in.w0);
  if (in.sphere1) g_l1 += cross(in.w1, g_tv);
  // n = -u (rounded components carry no derivative)
  g_u = d3{w.nmask[0] ? -g_n.x : 0., w.nmask[1] ? -g_n.y : 0., w.nmask[2] ? -g_n.z : 0.};
  // ov = d u
  g_d = dot(g_ov, u);
  g_u += in.d * g_ov;
  // l0 = (r0 + kappa d) u
  g_d += in.kappa * dot(g_l0, u);
  g_u += (in.r0 + in.kappa * in.d) * g_l0;
  // l1 = -(r1 + d/2) u
  if (in.sphere1) {
    g_d += -0.5 * dot(g_l1, u);
    g_u += (-(in.r1 + 0.5 * in.d)) * g_l1;
  }
  g.x0 = d3{0., 0., 0.};
}

// Sphere-sphere: u = AB/|AB|, d = |AB| - (r0 + r1), AB = x1 - x0.
DEM_HD void pair_vjp(d3 AB, ContactIn in, d3 F0b, d3 T0b, d3 F1b, d3 T1b, PairGrad& g) {
  double nab = norm(AB);
  in.u = AB / nab;
  in.d = nab - (in.r0 + in.r1);
  in.kappa = 0.5;
  in.sphere1 = true;
  ContactFwd w;
  contact_forward(in, w);
  d3 g_u;
  double g_d;
  contact_backward(in, w, F0b, T0b, F1b, T1b, g, g_u, g_d);
  // u = AB / nab, nab = |AB|, d = nab - R
  d3 g_AB = (g_u - dot(g_u, in.u) * in.u) / nab;
  g_AB += g_d * in.u;
  g.x0 = -g_AB;
}

// Sphere-box: with the centre in box coordinates c = Binv x0 + o, the
// closest feature is found by clamping (mask m = components outside the
// extents: one -> face, two -> edge, three -> corner); D = B (c - clamp(c))
// points from the feature to the centre, u = -D/|D|, d = |D| - r0; dD/dx0 =
// B diag(m) Binv.
DEM_HD void wall_vjp(const double* B, const double* Binv, const double* invorigin, const double* extent, d3 x0,
                     ContactIn in, d3 F0b, d3 T0b, PairGrad& g) {
  d3 c{dot(d3{Binv[0], Binv[1], Binv[2]}, x0) + invorigin[0], dot(d3{Binv[3], Binv[4], Binv[5]}, x0) + invorigin[1],
       dot(d3{Binv[6], Binv[7], Binv[8]}, x0) + invorigin[2]};
  double cl[3] = {c.x, c.y, c.z}, dl[3];
  int mask[3], nout = 0;
  for (int a = 0; a < 3; ++a) {
    double e = extent[a];
    double q = cl[a] > e ? e : (cl[a] < -e ? -e : cl[a]);
    mask[a] = (cl[a] > e || cl[a] < -e) ? 1 : 0;
    nout += mask[a];
    dl[a] = cl[a] - q;
  }
  d3 D{B[0] * dl[0] + B[1] * dl[1] + B[2] * dl[2], B[3] * dl[0] + B[4] * dl[1] + B[5] * dl[2],
       B[6] * dl[0] + B[7] * dl[1] + B[8] * dl[2]};
  double nd = norm(D);
  in.u = (-1. / nd) * D;
  in.d = nd - in.r0;
  in.kappa = nout <= 1 ? 0.5 : 1.5;
  in.sphere1 = false;
  ContactFwd w;
  contact_forward(in, w);
  d3 g_u;
  double g_d;
  contact_backward(in, w, F0b, T0b, d3{0., 0., 0.}, d3{0., 0., 0.}, g, g_u, g_d);
  // u = -D/nd, d = nd - r0
  d3 g_D = (-1. / nd) * (g_u - dot(g_u, in.u) * in.u);
  g_D += (-g_d) * in.u;
  // x0-gradient: B diag(mask) Binv g_D
  double t[3] = {dot(d3{Binv[0], Binv[1], Binv[2]}, g_D), dot(d3{Binv[3], Binv[4], Binv[5]}, g_D),
                 dot(d3{Binv[6], Binv[7], Binv[8]}, g_D)};
  for (int a = 0; a < 3; ++a) t[a] *= mask[a];
  g.x0 = d3{B[0] * t[0] + B[1] * t[1] + B[2] * t[2], B[3] * t[0] + B[4] * t[1] + B[5] * t[2],
            B[6] * t[0] + B[7] * t[1] + B[8] * t[2]};
}

// Force on body 0 of one contact in exact arithmetic (used by the gradient
// checks of the derivative code)
DEM_HD d3 contact_force(const ContactIn& in) {
  ContactFwd w;
  contact_forward(in, w);
  return w.f;
}

}  // namespace dem
