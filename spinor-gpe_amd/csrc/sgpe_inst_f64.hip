// complex128 instantiations of the split-step, FFT and energy kernels.
#include "sgpe_registry.h"

SGPE_DEFINE_REGISTRY(registry_f64, double, 0, SGPE_DEF_COLS_F64, SGPE_COL_SHAPES_F64_ONLY)
