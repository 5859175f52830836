// Helmholtz matrix-core M2L, M2M and L2L kernels of orders 26..27 (FP64
// and FP32), instantiated from hel_m2l_impl.h.

#include "hel_m2l_impl.h"

namespace wbfmm {
WBFMM_HEL_RANGE_DEFINE(26, 27)
}  // namespace wbfmm
