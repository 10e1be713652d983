// special_functions.h -- the reference header of this name (include/special_functions.h), served by the MI355X class
// layer declared in nrcps_host.h.
#include "nrcps_host.h"
