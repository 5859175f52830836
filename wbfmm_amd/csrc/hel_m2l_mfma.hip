// hel_m2l_mfma.hip -- Helmholtz M2L as FP64 matrix-core products (gfx950).
//
// The reference translates each source expansion x by rotating it into the
// frame of the offset (H(theta), phases e^{i m chi} in and e^{-i nu phi}
// out), shifting it along the axis with the complex coaxial S|R table, and
// rotating back (phases e^{i m phi} in, e^{-i nu chi} out)
// (src/pass.c:56-105).  The phi phases of the two rotations meet across the
// coaxial stage, which keeps the order m, and cancel.  With the
// phase-rotated x'_m = e^{i m chi} x_m and
//   P_m = x'_m + x'_{-m},  Q_m = x'_m - x'_{-m}   (m >= 0, P_0 = 2 x'_0)
// the real H(theta) of degree n, for which h(-nu,-m) = h(nu,m), splits into
//   a cos block  A_{nu m} = h(nu,m) + h(-nu,m) (m > 0), A_{nu 0} = h(nu,0)
//   on P (nu, m = 0..n), and a sin block B_{nu m} = h(nu,m) - h(-nu,m) on
//   Q (nu, m = 1..n);
// the coaxial table of order m acts alike on P and Q; and after the second
// rotation (the same tiles) y_nu = (P + Q)/2, y_{-nu} = (P - Q)/2 and
// L_n^nu += e^{-i nu chi} y_nu.  These are the block shapes of the Laplace
// M2L, so the Laplace plan of the same order (bins, slot maps, per-offset
// theta and distance indices, the cos/sin of m chi) carries over: a block
// of W = 4 (orders below 16) or 8 wavefronts takes 8 target boxes with the
// same Morton child digit,
// and the 16 MFMA columns are (target t, part s) = 2t + s with s = 0 the
// real and s = 1 the imaginary part.  The rotation tiles are real and act
// on both parts at once; the coaxial stage takes two products per tile step,
// the real part of the table on the column and the imaginary part on the
// partner column (col ^ 1).  Wavefront w computes the bins b = w (mod W) of
// every stage; its lane (col, grp) owns the units u = grp + 4 w + 4 W i: for each it loads x_{+m} (s = 0) or x_{-m} (s = 1),
// applies its phase, exchanges the result with its partner lane and writes
// the real (s = 0) or imaginary (s = 1) parts of P and Q; after the third
// stage it reads both parts of P and Q and accumulates L^{+nu} (s = 0) or
// L^{-nu} (s = 1).  Expansions are box-major, (N+1)^2 complex coefficients
// per box at n(n+1)+m.  Every LDS slot has one writer and the sums run in
// a fixed order: results do not depend on the launch.

#include <cstdio>

#include <hip/hip_runtime.h>

#include "hip_common.h"
#include "m2l_mfma.h"
#include "hel_m2l_ranges.h"

namespace wbfmm {

template <typename T>
static int hel_m2l_entry(const uint32_t *cls, const unsigned *cstart, const unsigned *ccount,
                         const uint32_t *sstart, const uint32_t *send, unsigned level, int N,
                         const T *mps, T *mpr, const void *plan_dev, const T *A, const T *A2,
                         const T *CS, const wbfmm_m2l_plan_t *plan_host) {
  ClassArgs ca;
  ca.qc = 0;
  unsigned nch = 0;
  for (int c = 0; c < 8; c++) {
    ca.start[c] = cstart[c];
    ca.count[c] = ccount[c];
    ca.chunk_prefix[c] = nch;
    nch += (ccount[c] + 7) / 8;
  }
  ca.chunk_prefix[8] = nch;
  if (nch == 0) return 0;
  // source boxes are found by 30-bit Morton arithmetic (levels up to 10;
  // a dense level array above that exceeds any device memory)
  if (level < 2 || level > (unsigned)kMortonAddMaxLevel) return WBFMM_HIP_ERR_ARG;
  if (N < 2 || N > 27) return WBFMM_HIP_ERR_ORDER;
  if (!plan_host || plan_host->N != N) return WBFMM_HIP_ERR_ARG;
  {
    const int rz = ensure_zero();
    if (rz) return rz;
  }
  uint32_t *occ = nullptr;
  {
    const int ro = occupancy_bits(sstart, send, level, &occ);
    if (ro) return ro;
  }
  const T *zero = (const T *)zero_expansion();
  const int rc = hel_m2l_orders<T>(N, cls, ca, nch, occ, (int)level, mps, mpr, plan_dev, A, A2,
                                   CS, zero, plan_host);
  dfree(occ);
  return rc;
}

template <typename T>
static int hel_shift_entry(int mode, const uint32_t *parents, unsigned npar,
                           const uint32_t *cstart, const uint32_t *cend, const uint32_t *cflag,
                           int N, const T *in, T *out, const void *plan_dev, const T *A,
                           const T *A2, const T *CS, const wbfmm_m2l_plan_t *plan_host) {
  if (npar == 0) return 0;
  if (mode != 0 && mode != 1) return WBFMM_HIP_ERR_ARG;
  if (mode == 0 && (!cstart || !cend)) return WBFMM_HIP_ERR_ARG;
  if (N < 2 || N > 27) return WBFMM_HIP_ERR_ORDER;
  if (!plan_host || plan_host->N != N) return WBFMM_HIP_ERR_ARG;
  {
    const int rz = ensure_zero();
    if (rz) return rz;
  }
  const T *zero = (const T *)zero_expansion();
  return hel_shift_orders<T>(mode, N, parents, npar, cstart, cend, cflag, in, out, plan_dev, A,
                             A2, CS, zero, plan_host);
}

}  // namespace wbfmm

using namespace wbfmm;

extern "C" {

int wbfmm_hip_hel_m2l_mfma(const uint32_t *cls, const unsigned *cstart, const unsigned *ccount,
                           const uint32_t *sstart, const uint32_t *send, unsigned level, int N,
                           const double *mps, double *mpr, const void *plan_dev,
                           const double *A, const double *A2, const double *CS,
                           const wbfmm_m2l_plan_t *plan_host) {
  return hel_m2l_entry<double>(cls, cstart, ccount, sstart, send, level, N, mps, mpr, plan_dev,
                               A, A2, CS, plan_host);
}

int wbfmm_hip_hel_m2l_mfma_f32(const uint32_t *cls, const unsigned *cstart,
                               const unsigned *ccount, const uint32_t *sstart,
                               const uint32_t *send, unsigned level, int N, const float *mps,
                               float *mpr, const void *plan_dev, const float *A, const float *A2,
                               const float *CS, const wbfmm_m2l_plan_t *plan_host) {
  return hel_m2l_entry<float>(cls, cstart, ccount, sstart, send, level, N, mps, mpr, plan_dev,
                              A, A2, CS, plan_host);
}

int wbfmm_hip_hel_shift_mfma(int mode, const uint32_t *parents, unsigned npar,
                             const uint32_t *cstart, const uint32_t *cend,
                             const uint32_t *cflag, int N, const double *in, double *out,
                             const void *plan_dev, const double *A, const double *A2,
                             const double *CS, const wbfmm_m2l_plan_t *plan_host) {
  return hel_shift_entry<double>(mode, parents, npar, cstart, cend, cflag, N, in, out, plan_dev,
                                 A, A2, CS, plan_host);
}

int wbfmm_hip_hel_shift_mfma_f32(int mode, const uint32_t *parents, unsigned npar,
                                 const uint32_t *cstart, const uint32_t *cend,
                                 const uint32_t *cflag, int N, const float *in, float *out,
                                 const void *plan_dev, const float *A, const float *A2,
                                 const float *CS, const wbfmm_m2l_plan_t *plan_host) {
  return hel_shift_entry<float>(mode, parents, npar, cstart, cend, cflag, N, in, out, plan_dev,
                                A, A2, CS, plan_host);
}

}  // extern "C"
