// overlap_mfma.hip -- split-BF16x3 overlap products on MFMA (placeholder unit).
#include "nrc_device.h"
#include "nrc_internal.h"
