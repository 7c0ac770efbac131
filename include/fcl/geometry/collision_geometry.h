// Forwarding header at FCL's include path (include/fcl/geometry/collision_geometry.h):
// everything this path needs is declared in fcl/fcl.h.
#ifndef FCL_AMD_FWD_GEOMETRY_COLLISION_GEOMETRY_H
#define FCL_AMD_FWD_GEOMETRY_COLLISION_GEOMETRY_H
#include "fcl/fcl.h"
#endif
