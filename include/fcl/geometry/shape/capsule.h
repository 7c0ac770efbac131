// Forwarding header at FCL's include path (include/fcl/geometry/shape/capsule.h):
// everything this path needs is declared in fcl/fcl.h.
#ifndef FCL_AMD_FWD_GEOMETRY_SHAPE_CAPSULE_H
#define FCL_AMD_FWD_GEOMETRY_SHAPE_CAPSULE_H
#include "fcl/fcl.h"
#endif
