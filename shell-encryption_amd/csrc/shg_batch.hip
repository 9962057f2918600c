// Persistent launch over uneven limb jobs (filled in with the batching work).
#include <hip/hip_runtime.h>

#include "shg_device.h"
#include "shg_kernels.h"
