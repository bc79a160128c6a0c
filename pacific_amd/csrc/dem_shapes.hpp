// Host construction of convex shape classes (see dem_shapes.cpp).
#pragma once
#include <string>
#include <vector>

#include "dem_types.hpp"

namespace dem {

struct ShapeBuild {
  ShapeDesc desc;
  double volume;             // per unit density
  double inertia[6];         // xx, xy, xz, yy, yz, zz per unit density
  std::vector<double> verts; // polyhedron vertices (3 per vertex)
};

// type: kShapeSphere (radius), kShapeBox (edge lengths L[3]) or kShapePoly
// (nv vertices, nf faces of face_sizes[f] indices each).  Returns 0, or 1
// with a message in err.
int build_shape(int type, double radius, const double* L, int nv, const double* verts, int nf,
                const int* face_sizes, const int* face_idx, double crust, ShapeBuild& out, std::string& err);

// quaternion (x, y, z, w) of a row-major rotation matrix
void quat_from_basis(const double* B, double* q);

}  // namespace dem
