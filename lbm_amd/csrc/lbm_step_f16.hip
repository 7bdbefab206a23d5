// Step kernels for __half population storage.
#include "lbm_step_kernels.h"

LBM_INSTANTIATE_STEPS(__half)
