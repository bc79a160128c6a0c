// Force sweep launchers: spheres, FP64.
#define DEM_FORCES_MIXED 0
#define DEM_FORCES_SHAPES 0
#include "dem_forces_impl.hpp"
