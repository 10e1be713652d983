// global_job_parameter.h -- the reference header of this name (include/global_job_parameter.h), served by the MI355X class
// layer declared in nrcps_host.h.
#include "nrcps_host.h"
