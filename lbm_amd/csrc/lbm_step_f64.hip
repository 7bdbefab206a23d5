// Step kernels for double population storage.
#include "lbm_step_kernels.h"

LBM_INSTANTIATE_STEPS(double)
