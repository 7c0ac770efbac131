// Forwarding header at FCL's include path (include/fcl/narrowphase/gjk_solver_type.h):
// everything this path needs is declared in fcl/fcl.h.
#ifndef FCL_AMD_FWD_NARROWPHASE_GJK_SOLVER_TYPE_H
#define FCL_AMD_FWD_NARROWPHASE_GJK_SOLVER_TYPE_H
#include "fcl/fcl.h"
#endif
