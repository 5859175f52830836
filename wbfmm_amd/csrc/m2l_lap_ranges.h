// m2l_lap_ranges.h -- the order ranges of the Laplace matrix-core
// translation kernels.  Each range is instantiated in a translation unit
// of its own (m2l_lap_o<LO>_<HI>.hip, from m2l_lap_impl.h) so that the
// orders compile in parallel; m2l_mfma.hip calls lap_m2l_orders /
// lap_shift_orders, which pick the range of an order.

#ifndef WBFMM_M2L_LAP_RANGES_H
#define WBFMM_M2L_LAP_RANGES_H

#include <stdint.h>

#include "m2l_mfma.h"
#include "wbfmm_hip.h"

namespace wbfmm {

// blocks per chunk of the split M2L form: levels with fewer than 1024
// chunks spread each chunk's offsets over up to 8 blocks (2048 in all)
static inline int m2l_nsplit(unsigned nchunks) {
  return nchunks >= 1024 ? 1 : (int)(2048 / nchunks < 8 ? 2048 / nchunks : 8);
}

#define WBFMM_LAP_RANGES(X) X(2, 7) X(8, 10) X(11, 12) X(13, 15) X(16, 18) X(19, 21) X(22, 24)

#define WBFMM_LAP_RANGE_DECLARE(LO, HI)                                                      \
  int lap_m2l_range_##LO##_##HI(int N, const uint32_t *cls, const ClassArgs &ca,            \
                                unsigned nch, const uint32_t *occ, int level,               \
                                const double *mps, double *mpr, const void *plan,           \
                                const double *A1, const double *A3, const double *A2,       \
                                const double *CS, const double *zero,                       \
                                const wbfmm_m2l_plan_t *hplan, double *part);               \
  int lap_m2l_range_##LO##_##HI(int N, const uint32_t *cls, const ClassArgs &ca,            \
                                unsigned nch, const uint32_t *occ, int level,               \
                                const float *mps, float *mpr, const void *plan,             \
                                const float *A1, const float *A3, const float *A2,          \
                                const float *CS, const float *zero,                         \
                                const wbfmm_m2l_plan_t *hplan, float *part);                \
  int lap_shift_range_##LO##_##HI(int mode, int N, const uint32_t *parents, unsigned npar,  \
                                  const uint32_t *cstart, const uint32_t *cend,             \
                                  const uint32_t *cflag, const double *in, double *out,     \
                                  const void *plan, const double *A1, const double *A3,     \
                                  const double *A2, const double *CS, const double *zero,   \
                                  const wbfmm_m2l_plan_t *hplan);                           \
  int lap_shift_range_##LO##_##HI(int mode, int N, const uint32_t *parents, unsigned npar,  \
                                  const uint32_t *cstart, const uint32_t *cend,             \
                                  const uint32_t *cflag, const float *in, float *out,       \
                                  const void *plan, const float *A1, const float *A3,       \
                                  const float *A2, const float *CS, const float *zero,      \
                                  const wbfmm_m2l_plan_t *hplan);
WBFMM_LAP_RANGES(WBFMM_LAP_RANGE_DECLARE)
#undef WBFMM_LAP_RANGE_DECLARE

// the M2L of order N (T = double or float)
template <typename T>
inline int lap_m2l_orders(int N, const uint32_t *cls, const ClassArgs &ca, unsigned nch,
                          const uint32_t *occ, int level, const T *mps, T *mpr,
                          const void *plan, const T *A1, const T *A3, const T *A2,
                          const T *CS, const T *zero, const wbfmm_m2l_plan_t *hplan,
                          T *part) {
#define WBFMM_LAP_RANGE_CALL(LO, HI)                                                     \
  if (N >= LO && N <= HI)                                                               \
    return lap_m2l_range_##LO##_##HI(N, cls, ca, nch, occ, level, mps, mpr, plan, A1, A3, \
                                     A2, CS, zero, hplan, part);
  WBFMM_LAP_RANGES(WBFMM_LAP_RANGE_CALL)
#undef WBFMM_LAP_RANGE_CALL
  return WBFMM_HIP_ERR_ORDER;
}

// the M2M (mode 0) or L2L (mode 1) of order N
template <typename T>
inline int lap_shift_orders(int mode, int N, const uint32_t *parents, unsigned npar,
                            const uint32_t *cstart, const uint32_t *cend, const uint32_t *cflag,
                            const T *in, T *out, const void *plan, const T *A1, const T *A3,
                            const T *A2, const T *CS, const T *zero,
                            const wbfmm_m2l_plan_t *hplan) {
#define WBFMM_LAP_RANGE_CALL(LO, HI)                                                       \
  if (N >= LO && N <= HI)                                                                 \
    return lap_shift_range_##LO##_##HI(mode, N, parents, npar, cstart, cend, cflag, in, out, \
                                       plan, A1, A3, A2, CS, zero, hplan);
  WBFMM_LAP_RANGES(WBFMM_LAP_RANGE_CALL)
#undef WBFMM_LAP_RANGE_CALL
  return WBFMM_HIP_ERR_ORDER;
}

}  // namespace wbfmm

#endif
