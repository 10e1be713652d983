// two_body.h -- the reference header of this name (include/two_body.h), served by the MI355X class
// layer declared in nrcps_host.h.
#include "nrcps_host.h"
