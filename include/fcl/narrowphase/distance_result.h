// Forwarding header at FCL's include path (include/fcl/narrowphase/distance_result.h):
// everything this path needs is declared in fcl/fcl.h.
#ifndef FCL_AMD_FWD_NARROWPHASE_DISTANCE_RESULT_H
#define FCL_AMD_FWD_NARROWPHASE_DISTANCE_RESULT_H
#include "fcl/fcl.h"
#endif
