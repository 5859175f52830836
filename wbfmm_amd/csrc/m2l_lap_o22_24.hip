// Laplace matrix-core M2L, M2M and L2L kernels of orders 22..24 (FP64 and
// FP32), instantiated from m2l_lap_impl.h.

#include "m2l_lap_impl.h"

namespace wbfmm {
WBFMM_LAP_RANGE_DEFINE(22, 24)
}  // namespace wbfmm
