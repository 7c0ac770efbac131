// Forwarding header at FCL's include path (include/fcl/math/bv/AABB.h):
// everything this path needs is declared in fcl/fcl.h.
#ifndef FCL_AMD_FWD_MATH_BV_AABB_H
#define FCL_AMD_FWD_MATH_BV_AABB_H
#include "fcl/fcl.h"
#endif
