// Step kernels for float population storage.
#include "lbm_step_kernels.h"

LBM_INSTANTIATE_STEPS(float)
