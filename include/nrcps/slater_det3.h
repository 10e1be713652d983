// slater_det3.h -- the reference header of this name (include/slater_det3.h), served by the MI355X class
// layer declared in nrcps_host.h.
#include "nrcps_host.h"
