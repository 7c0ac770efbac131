// Forwarding header at FCL's include path (include/fcl/broadphase/broadphase_dynamic_AABB_tree.h):
// everything this path needs is declared in fcl/fcl.h.
#ifndef FCL_AMD_FWD_BROADPHASE_DYNAMIC_AABB_TREE_H
#define FCL_AMD_FWD_BROADPHASE_DYNAMIC_AABB_TREE_H
#include "fcl/fcl.h"
#endif
