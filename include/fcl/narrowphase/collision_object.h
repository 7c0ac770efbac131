// Forwarding header at FCL's include path (include/fcl/narrowphase/collision_object.h):
// everything this path needs is declared in fcl/fcl.h.
#ifndef FCL_AMD_FWD_NARROWPHASE_COLLISION_OBJECT_H
#define FCL_AMD_FWD_NARROWPHASE_COLLISION_OBJECT_H
#include "fcl/fcl.h"
#endif
