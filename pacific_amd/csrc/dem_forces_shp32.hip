// Force sweep launchers: non-spherical bodies, mixed precision.
#define DEM_FORCES_MIXED 1
#define DEM_FORCES_SHAPES 1
#include "dem_forces_impl.hpp"
