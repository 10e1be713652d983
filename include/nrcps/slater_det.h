// slater_det.h -- the reference header of this name (include/slater_det.h), served by the MI355X class
// layer declared in nrcps_host.h.
#include "nrcps_host.h"
