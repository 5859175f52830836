// Helmholtz matrix-core M2L, M2M and L2L kernels of orders 19..21 (FP64
// and FP32), instantiated from hel_m2l_impl.h.

#include "hel_m2l_impl.h"

namespace wbfmm {
WBFMM_HEL_RANGE_DEFINE(19, 21)
}  // namespace wbfmm
