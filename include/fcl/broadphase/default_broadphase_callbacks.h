// Forwarding header at FCL's include path (include/fcl/broadphase/default_broadphase_callbacks.h):
// everything this path needs is declared in fcl/fcl.h.
#ifndef FCL_AMD_FWD_DEFAULT_BROADPHASE_CALLBACKS_H
#define FCL_AMD_FWD_DEFAULT_BROADPHASE_CALLBACKS_H
#include "fcl/fcl.h"
#endif
