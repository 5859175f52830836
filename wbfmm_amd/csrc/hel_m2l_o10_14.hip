// Helmholtz matrix-core M2L, M2M and L2L kernels of orders 10..14 (FP64
// and FP32), instantiated from hel_m2l_impl.h.

#include "hel_m2l_impl.h"

namespace wbfmm {
WBFMM_HEL_RANGE_DEFINE(10, 14)
}  // namespace wbfmm
