// Forwarding header at FCL's include path (include/fcl/geometry/shape/box.h):
// everything this path needs is declared in fcl/fcl.h.
#ifndef FCL_AMD_FWD_GEOMETRY_SHAPE_BOX_H
#define FCL_AMD_FWD_GEOMETRY_SHAPE_BOX_H
#include "fcl/fcl.h"
#endif
