// constants.h -- the reference header of this name (include/constants.h), served by the MI355X class
// layer declared in nrcps_host.h.
#include "nrcps_host.h"
