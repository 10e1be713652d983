// slater_det2.h -- the reference header of this name (include/slater_det2.h), served by the MI355X class
// layer declared in nrcps_host.h.
#include "nrcps_host.h"
