// Forwarding header at FCL's include path (include/fcl/geometry/bvh/BVH_internal.h):
// everything this path needs is declared in fcl/fcl.h.
#ifndef FCL_AMD_FWD_GEOMETRY_BVH_BVH_INTERNAL_H
#define FCL_AMD_FWD_GEOMETRY_BVH_BVH_INTERNAL_H
#include "fcl/fcl.h"
#endif
