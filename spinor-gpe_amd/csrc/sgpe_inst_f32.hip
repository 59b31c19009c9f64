// complex64 instantiations of the split-step, FFT and energy kernels.
#include "sgpe_registry.h"

SGPE_DEFINE_REGISTRY(registry_f32, float, 1, SGPE_DEF_COLS_F32, SGPE_COL_SHAPES_F32_ONLY)
