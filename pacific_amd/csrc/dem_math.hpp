// FP64 three-vector and quaternion primitives evaluated in the exact
// operation order of Grains3D's Group3/Vector3/Quaternion classes, so that a
// kernel built with -ffp-contract=off reproduces the CPU results bit for bit.
//   Group3 arithmetic ............ GRAINS_V1/Grains/Geometry/src/Group3.cpp:138-320
//   cross product, Norm .......... GRAINS_V1/Grains/Geometry/src/Vector3.cpp:114-152
//   round ........................ GRAINS_V1/Grains/Geometry/src/Group3.cpp:390-395
//   quaternion products .......... GRAINS_V1/Grains/Geometry/src/Quaternion.cpp:74-115, 442-470
#pragma once
#include <hip/hip_runtime.h>
#include <cmath>

#define DEM_HD __host__ __device__ __forceinline__

namespace dem {

constexpr double kPi = 3.14159265358979323846;   // Basic.hh:22
constexpr double kEps = 1.0e-10;                 // Basic.hh:26 (EPSILON)
constexpr double kObstacleMass = 1.e20;          // Component.cpp:16

struct d3 {
  double x, y, z;
};

DEM_HD d3 mk(double x, double y, double z) { return d3{x, y, z}; }
DEM_HD d3 operator+(d3 a, d3 b) { return d3{a.x + b.x, a.y + b.y, a.z + b.z}; }
DEM_HD d3 operator-(d3 a, d3 b) { return d3{a.x - b.x, a.y - b.y, a.z - b.z}; }
DEM_HD d3 operator-(d3 a) { return d3{-a.x, -a.y, -a.z}; }
// Group3 * scalar and scalar * Group3 both scale component-wise (comp * d)
DEM_HD d3 operator*(d3 a, double d) { return d3{a.x * d, a.y * d, a.z * d}; }
DEM_HD d3 operator*(double d, d3 a) { return d3{a.x * d, a.y * d, a.z * d}; }
DEM_HD d3 operator/(d3 a, double d) { return d3{a.x / d, a.y / d, a.z / d}; }
DEM_HD d3& operator+=(d3& a, d3 b) { a.x += b.x; a.y += b.y; a.z += b.z; return a; }
DEM_HD d3& operator-=(d3& a, d3 b) { a.x -= b.x; a.y -= b.y; a.z -= b.z; return a; }
DEM_HD d3& operator*=(d3& a, double d) { a.x *= d; a.y *= d; a.z *= d; return a; }
DEM_HD d3& operator/=(d3& a, double d) { a.x /= d; a.y /= d; a.z /= d; return a; }
DEM_HD double dot(d3 a, d3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
DEM_HD d3 cross(d3 a, d3 b) {
  return d3{a.y * b.z - a.z * b.y, -a.x * b.z + a.z * b.x, a.x * b.y - a.y * b.x};
}
DEM_HD double norm(d3 a) { return sqrt(a.x * a.x + a.y * a.y + a.z * a.z); }
DEM_HD void round_small(d3& a) {
  a.x = fabs(a.x) < kEps ? 0. : a.x;
  a.y = fabs(a.y) < kEps ? 0. : a.y;
  a.z = fabs(a.z) < kEps ? 0. : a.z;
}

// Quaternion [w, v]
struct quat {
  d3 v;
  double w;
};
DEM_HD quat conj(quat q) { return quat{-q.v, q.w}; }
// this x rhs
DEM_HD quat qmul(quat a, quat b) {
  double tmp = (a.w * b.w) - dot(a.v, b.v);
  d3 vt = (cross(a.v, b.v) + (a.w * b.v)) + (b.w * a.v);
  return quat{vt, tmp};
}
// [0, lhs] x q  (the (Vector3 , Quaternion) operator)
DEM_HD quat vec_q(d3 lhs, quat q) {
  double tmp = dot(-lhs, q.v);
  d3 vt = cross(lhs, q.v) + (lhs * q.w);
  return quat{vt, tmp};
}
// vector part of this x rhs
DEM_HD d3 mult_to_vec(quat a, quat b) { return (cross(a.v, b.v) + (a.w * b.v)) + (b.w * a.v); }
// body-fixed coordinates of a space vector: vec( conj(q) x ([0,u] x q) )
DEM_HD d3 to_body(quat q, d3 u) { return mult_to_vec(conj(q), vec_q(u, q)); }
// space coordinates of a body vector: vec( q x ([0,u] x conj(q)) )
DEM_HD d3 to_space(quat q, d3 u) { return mult_to_vec(q, vec_q(u, conj(q))); }

}  // namespace dem
