// Step kernels for __hip_bfloat16 population storage.
#include "lbm_step_kernels.h"

LBM_INSTANTIATE_STEPS(__hip_bfloat16)
