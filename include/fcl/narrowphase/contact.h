// Forwarding header at FCL's include path (include/fcl/narrowphase/contact.h):
// everything this path needs is declared in fcl/fcl.h.
#ifndef FCL_AMD_FWD_NARROWPHASE_CONTACT_H
#define FCL_AMD_FWD_NARROWPHASE_CONTACT_H
#include "fcl/fcl.h"
#endif
