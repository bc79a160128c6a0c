// Force sweep launchers: non-spherical bodies, FP64.
#define DEM_FORCES_MIXED 0
#define DEM_FORCES_SHAPES 1
#include "dem_forces_impl.hpp"
