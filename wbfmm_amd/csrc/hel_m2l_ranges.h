// hel_m2l_ranges.h -- the order ranges of the Helmholtz matrix-core
// translation kernels.  Each range is instantiated in a translation unit
// of its own (hel_m2l_o<LO>_<HI>.hip, from hel_m2l_impl.h);
// hel_m2l_mfma.hip calls hel_m2l_orders / hel_shift_orders, which pick the
// range of an order.

#ifndef WBFMM_HEL_M2L_RANGES_H
#define WBFMM_HEL_M2L_RANGES_H

#include <stdint.h>

#include "m2l_mfma.h"
#include "wbfmm_hip.h"

namespace wbfmm {

#define WBFMM_HEL_RANGES(X) X(2, 9) X(10, 14) X(15, 18) X(19, 21) X(22, 23) X(24, 25) X(26, 27)

#define WBFMM_HEL_RANGE_DECLARE_T(LO, HI, T)                                                  \
  int hel_m2l_range_##LO##_##HI(int N, const uint32_t *cls, const ClassArgs &ca,             \
                                unsigned nch, const uint32_t *occ, int level, const T *mps,  \
                                T *mpr, const void *plan, const T *A, const T *A2,           \
                                const T *CS, const T *zero, const wbfmm_m2l_plan_t *hplan);  \
  int hel_shift_range_##LO##_##HI(int mode, int N, const uint32_t *parents, unsigned npar,   \
                                  const uint32_t *cstart, const uint32_t *cend,              \
                                  const uint32_t *cflag, const T *in, T *out,                \
                                  const void *plan, const T *A, const T *A2, const T *CS,    \
                                  const T *zero, const wbfmm_m2l_plan_t *hplan);
#define WBFMM_HEL_RANGE_DECLARE(LO, HI)     \
  WBFMM_HEL_RANGE_DECLARE_T(LO, HI, double) \
  WBFMM_HEL_RANGE_DECLARE_T(LO, HI, float)
WBFMM_HEL_RANGES(WBFMM_HEL_RANGE_DECLARE)
#undef WBFMM_HEL_RANGE_DECLARE
#undef WBFMM_HEL_RANGE_DECLARE_T

// the Helmholtz M2L of order N (T = double or float)
template <typename T>
inline int hel_m2l_orders(int N, const uint32_t *cls, const ClassArgs &ca, unsigned nch,
                          const uint32_t *occ, int level, const T *mps, T *mpr,
                          const void *plan, const T *A, const T *A2, const T *CS,
                          const T *zero, const wbfmm_m2l_plan_t *hplan) {
#define WBFMM_HEL_RANGE_CALL(LO, HI)                                                        \
  if (N >= LO && N <= HI)                                                                  \
    return hel_m2l_range_##LO##_##HI(N, cls, ca, nch, occ, level, mps, mpr, plan, A, A2, CS, \
                                     zero, hplan);
  WBFMM_HEL_RANGES(WBFMM_HEL_RANGE_CALL)
#undef WBFMM_HEL_RANGE_CALL
  return WBFMM_HIP_ERR_ORDER;
}

// the Helmholtz M2M (mode 0) or L2L (mode 1) of order N
template <typename T>
inline int hel_shift_orders(int mode, int N, const uint32_t *parents, unsigned npar,
                            const uint32_t *cstart, const uint32_t *cend, const uint32_t *cflag,
                            const T *in, T *out, const void *plan, const T *A, const T *A2,
                            const T *CS, const T *zero, const wbfmm_m2l_plan_t *hplan) {
#define WBFMM_HEL_RANGE_CALL(LO, HI)                                                          \
  if (N >= LO && N <= HI)                                                                    \
    return hel_shift_range_##LO##_##HI(mode, N, parents, npar, cstart, cend, cflag, in, out, \
                                       plan, A, A2, CS, zero, hplan);
  WBFMM_HEL_RANGES(WBFMM_HEL_RANGE_CALL)
#undef WBFMM_HEL_RANGE_CALL
  return WBFMM_HIP_ERR_ORDER;
}

}  // namespace wbfmm

#endif
