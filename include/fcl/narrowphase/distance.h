// Forwarding header at FCL's include path (include/fcl/narrowphase/distance.h):
// everything this path needs is declared in fcl/fcl.h.
#ifndef FCL_AMD_FWD_NARROWPHASE_DISTANCE_H
#define FCL_AMD_FWD_NARROWPHASE_DISTANCE_H
#include "fcl/fcl.h"
#endif
