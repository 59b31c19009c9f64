// complex128 instantiations of the persistent step kernel.
#include "sgpe_persist.h"

SGPE_DEFINE_PERSIST_REGISTRY(persist_registry_f64, double, 0)
