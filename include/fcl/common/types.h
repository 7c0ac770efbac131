// Forwarding header at FCL's include path (include/fcl/common/types.h):
// everything this path needs is declared in fcl/fcl.h.
#ifndef FCL_AMD_FWD_COMMON_TYPES_H
#define FCL_AMD_FWD_COMMON_TYPES_H
#include "fcl/fcl.h"
#endif
