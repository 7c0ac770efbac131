// Forwarding header at FCL's include path (include/fcl/broadphase/broadphase_collision_manager.h):
// everything this path needs is declared in fcl/fcl.h.
#ifndef FCL_AMD_FWD_BROADPHASE_COLLISION_MANAGER_H
#define FCL_AMD_FWD_BROADPHASE_COLLISION_MANAGER_H
#include "fcl/fcl.h"
#endif
