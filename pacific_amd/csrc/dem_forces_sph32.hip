// Force sweep launchers: spheres, mixed precision.
#define DEM_FORCES_MIXED 1
#define DEM_FORCES_SHAPES 0
#include "dem_forces_impl.hpp"
