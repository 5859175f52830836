// m2l_mfma.hip -- Laplace M2L as FP64 matrix-core products (gfx950).
//
// The reference translates one source box at a time: rotate the multipole
// into the frame of the offset, coaxial S|R shift, rotate back and add into
// the target's local expansion (src/pass-laplace.c:35-105).  Here a
// wavefront takes 16 target boxes with the same Morton child digit (hence
// the same list-4 offsets) and, for each offset, applies the translation
// to the 16 source expansions at once:
//
//   Zin     rotate each (re, im) pair of order m by m chi   (vector units)
//   stage 1 Y = H1(theta) X      bin-packed cos/sin blocks  (MFMA)
//   stage 2 Z = S Y              coaxial S|R per order      (MFMA)
//             inputs scaled by t^-n, outputs by t^-(n+1)
//   stage 3 W = H3(theta) Z      (MFMA)
//   Zout    L += rotation of each (tr, ti) pair by -nu chi (vector units)
//
// The rotation matrices depend on the polar angle only (49 tables), the
// coaxial ones on the offset length (15 tables per box width, t^-k factors
// folded in), and all are read from L2; the per-offset azimuthal factors
// (cos, sin of m chi) come from a table.  The plan (bins, slot maps,
// per-offset theta/distance indices) is built on the host by
// wbfmm_hip_lap_m2l_tables.  One LDS buffer R of S slots x 16 columns
// carries the data between stages (slot s of column j at R[16 s + j]).
// Source expansions are box-major (box b at b*C); lane (j, g) handles the
// units u = g + 4 i of column j: it loads them for the next offset while
// the current one is computed (orders up to 12), rotates them by Zin into
// R, and accumulates the same units of its target after Zout.  The local
// expansions of the 16 targets stay in registers for all offsets and are
// added to mpr once at the end.
//
// MFMA operand layout (16x16x4 f64, checked by wbfmm_hip_mfma_layout_probe):
// lane l supplies A[l%16][l/16] and B[l/16][l%16]; the result holds
// D[l/16 + 4q][l%16], q = 0..3.

#include <cstdio>
#include <cstdlib>
#include <cstring>

#include <hip/hip_runtime.h>
#include <rocprim/rocprim.hpp>

#include <mutex>

#include "hip_common.h"
#include "m2l_mfma.h"
#include "m2l_lap_ranges.h"

namespace wbfmm {

// occupancy bitmap of a level: bit b set when box b holds points
__global__ void k_occ_bits(const uint32_t *start, const uint32_t *end,
                           uint64_t nbox, uint32_t *bits) {
  const uint64_t wd = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (wd * 32 >= nbox) return;
  uint32_t v = 0;
  for (int j = 0; j < 32; j++) {
    const uint64_t b = wd * 32 + j;
    if (b < nbox && end[b] > start[b]) v |= 1u << j;
  }
  bits[wd] = v;
}

// the occupancy bitmap of a level of 2^(3 level) boxes into stream-ordered
// memory that the caller releases with dfree
int occupancy_bits(const uint32_t *sstart, const uint32_t *send, unsigned level,
                   uint32_t **occ) {
  const uint64_t nbox = (uint64_t)1 << (3 * level);
  const size_t words = (size_t)((nbox + 31) / 32);
  *occ = nullptr;
  if (dmalloc((void **)occ, words * sizeof(uint32_t))) return WBFMM_HIP_ERR_NOMEM;
  WBFMM_LAUNCH("m2l_occ_bits", k_occ_bits, dim3((unsigned)((words + 255) / 256)), dim3(256), 0,
               sstart, send, nbox, *occ);
  return 0;
}

// one 16x16x4 product with known operands; out gets D in lane order
__global__ void k_mfma_layout_probe(double *out) {
  const int l = threadIdx.x;
  // A[i][k] = 1 + i + 100 k, B[k][j] = 1 + 3 k + 1000 j
  const int ai = l % 16, ak = l / 16, bk = l / 16, bj = l % 16;
  double a = 1.0 + ai + 100.0 * ak, bb = 1.0 + 3.0 * bk + 1000.0 * bj;
  d4 d = __builtin_amdgcn_mfma_f64_16x16x4f64(a, bb, d4{0.0, 0.0, 0.0, 0.0},
                                             0, 0, 0);
  for (int q = 0; q < 4; q++) out[4 * l + q] = d[q];
}

// the same product in FP32 (values are small integers, exact in FP32)
__global__ void k_mfma_layout_probe_f32(float *out) {
  const int l = threadIdx.x;
  const int ai = l % 16, ak = l / 16, bk = l / 16, bj = l % 16;
  float a = 1.0f + ai + 100.0f * ak, bb = 1.0f + 3.0f * bk + 1000.0f * bj;
  f4 d = __builtin_amdgcn_mfma_f32_16x16x4f32(a, bb, f4{0.0f, 0.0f, 0.0f, 0.0f},
                                             0, 0, 0);
  for (int q = 0; q < 4; q++) out[4 * l + q] = d[q];
}

// per-class counts of a box list: a histogram in LDS per block, then one
// global add per class and block (integer adds: the counts are exact)
__global__ void k_class_count(const uint32_t *list, unsigned n,
                              unsigned *counts) {
  __shared__ unsigned h[8];
  if (threadIdx.x < 8) h[threadIdx.x] = 0u;
  __syncthreads();
  for (unsigned i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    atomicAdd(&h[list[i] & 7], 1u);
  __syncthreads();
  if (threadIdx.x < 8 && h[threadIdx.x]) atomicAdd(&counts[threadIdx.x], h[threadIdx.x]);
}


// the zero expansion (its bytes are zero for either precision, real or
// complex coefficients)
static double *g_zero = nullptr;
static std::mutex g_zero_mu;

int ensure_zero() {
  std::lock_guard<std::mutex> lk(g_zero_mu);
  if (!g_zero) {
    const size_t bytes = sizeof(double) * 2 * (WBFMM_M2L_MAXN + 1) * (WBFMM_M2L_MAXN + 1);
    double *z = nullptr;
    WBFMM_CHECK(hipMalloc(&z, bytes));
    WBFMM_CHECK(hipMemsetAsync(z, 0, bytes, stream()));
    // other streams read it as soon as it is published
    WBFMM_CHECK(hipStreamSynchronize(stream()));
    g_zero = z;
  }
  return 0;
}

const double *zero_expansion() { return g_zero; }

template <typename T>
static int shift_entry(int mode, const uint32_t *parents, unsigned npar,
                       const uint32_t *cstart, const uint32_t *cend,
                       const uint32_t *cflag, int N, const T *in, T *out,
                       const void *plan_dev, const T *A1, const T *A3,
                       const T *A2, const T *CS,
                       const wbfmm_m2l_plan_t *plan_host) {
  if (npar == 0) return 0;
  if (mode != 0 && mode != 1) return WBFMM_HIP_ERR_ARG;
  if (mode == 0 && (!cstart || !cend)) return WBFMM_HIP_ERR_ARG;
  int rc = ensure_zero();
  if (rc) return rc;
  const T *zero = (const T *)g_zero;
  return lap_shift_orders<T>(mode, N, parents, npar, cstart, cend, cflag, in, out, plan_dev,
                             A1, A3, A2, CS, zero, plan_host);
}

// Scratch of the allocation-free M2L (wbfmm_hip_lap_m2l_mfma_noalloc):
// one file-scope device array per GPU, laid out as
//   [0, 16 KiB)            the zero expansion read in place of empty boxes,
//   [16 KiB, +2 MiB)       the occupancy bitmap of the source level
//                          (8^8 bits: levels up to 8),
//   [16 KiB + 2 MiB, end)  the partial expansions of the split form.
// Host code reaches it only through the address from hipGetSymbolAddress,
// taken once per device and kept in g_scratch_dev.
constexpr size_t kScratchBytes = (size_t)64 << 20;
constexpr size_t kScratchOcc = (size_t)16 << 10;
constexpr size_t kScratchPart = kScratchOcc + ((size_t)2 << 20);
constexpr unsigned kScratchMaxLevel = 8;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];
constexpr int kScratchDevices = 64;
static unsigned char *g_scratch_dev[kScratchDevices];
static std::mutex g_scratch_mu;

static int scratch_address(unsigned char **out) {
  int dev = 0;
  WBFMM_CHECK(hipGetDevice(&dev));
  if (dev < 0 || dev >= kScratchDevices)
    return fail_msg("M2L scratch: device index out of range", WBFMM_HIP_ERR_ARG);
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_dev[dev]) {
    void *ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr) return fail("hipGetSymbolAddress(scratch)", e);
    g_scratch_dev[dev] = (unsigned char *)ptr;
  }
  *out = g_scratch_dev[dev];
  return 0;
}

__global__ void k_fill_zero_u32(uint32_t *p, size_t n) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0u;
}

// ext: scratch of the allocation-free entry point (zero expansion,
// occupancy bits and partials at the offsets above), or null to allocate
template <typename T>
static int m2l_entry(const uint32_t *cls, const unsigned *cstart, const unsigned *ccount,
                     const uint32_t *sstart, const uint32_t *send, unsigned level, int N,
                     const T *mps, T *mpr, const void *plan_dev, const T *A1, const T *A3,
                     const T *A2, const T *CS, const wbfmm_m2l_plan_t *plan_host,
                     unsigned char *ext = nullptr) {
  ClassArgs ca;
  ca.qc = 0;
  unsigned nch = 0;
  for (int c = 0; c < 8; c++) {
    ca.start[c] = cstart[c];
    ca.count[c] = ccount[c];
    ca.chunk_prefix[c] = nch;
    nch += (ccount[c] + 15) / 16;
  }
  ca.chunk_prefix[8] = nch;
  if (nch == 0) return 0;
  // source boxes are found by 30-bit Morton arithmetic (levels up to 10;
  // a dense level array above that exceeds any device memory)
  if (level > (unsigned)kMortonAddMaxLevel) return WBFMM_HIP_ERR_ARG;
  if (N < 2 || N > 24) return WBFMM_HIP_ERR_ORDER;
  const uint64_t nbox = (uint64_t)1 << (3 * level);
  const size_t words = (size_t)((nbox + 31) / 32);
  const T *zero = nullptr;
  uint32_t *occ = nullptr;
  T *part = nullptr;
  if (ext) {
    // every size is checked before anything is launched
    char msg[256];
    if (level > kScratchMaxLevel) {
      snprintf(msg, sizeof(msg),
               "M2L without allocation: level %u above %u (the occupancy bitmap of "
               "8^level bits would exceed its %zu-byte scratch region)",
               level, kScratchMaxLevel, kScratchPart - kScratchOcc);
      return fail_msg(msg, WBFMM_HIP_ERR_ARG);
    }
    unsigned ntot = 0;
    for (int c = 0; c < 8; c++) ntot += ccount[c];
    const int nsplit = m2l_nsplit(nch);
    const size_t need =
        nsplit > 1 ? sizeof(T) * (size_t)nsplit * ntot * (size_t)((N + 1) * (N + 1)) : 0;
    if (need > kScratchBytes - kScratchPart) {
      snprintf(msg, sizeof(msg),
               "M2L without allocation: %u targets at order %d need %zu bytes of split "
               "partials, above the %zu-byte scratch region",
               ntot, N, need, kScratchBytes - kScratchPart);
      return fail_msg(msg, WBFMM_HIP_ERR_ARG);
    }
    // the zero expansion must read zero whatever an earlier call left
    const size_t zw = kScratchOcc / sizeof(uint32_t);
    WBFMM_LAUNCH("m2l_scratch_zero", k_fill_zero_u32, dim3((unsigned)((zw + 255) / 256)),
                 dim3(256), 0, (uint32_t *)ext, zw);
    zero = (const T *)ext;
    occ = (uint32_t *)(ext + kScratchOcc);
    part = nsplit > 1 ? (T *)(ext + kScratchPart) : nullptr;
  } else {
    const int rz = ensure_zero();
    if (rz) return rz;
    zero = (const T *)g_zero;
    // occupancy bits of the source level, per call (stream-ordered)
    if (dmalloc((void **)&occ, words * sizeof(uint32_t))) return WBFMM_HIP_ERR_NOMEM;
  }
  WBFMM_LAUNCH("m2l_occ_bits", k_occ_bits, dim3((unsigned)((words + 255) / 256)), dim3(256), 0,
               sstart, send, nbox, occ);
  const int rc = lap_m2l_orders<T>(N, cls, ca, nch, occ, (int)level, mps, mpr, plan_dev, A1, A3,
                                   A2, CS, zero, plan_host, part);
  if (!ext) dfree(occ);
  return rc;
}

}  // namespace wbfmm

using namespace wbfmm;

extern "C" {

int wbfmm_hip_split_classes(const uint32_t *list, unsigned n, uint32_t *out,
                            unsigned *cstart, unsigned *ccount) {
  for (int c = 0; c < 8; c++) cstart[c] = ccount[c] = 0;
  if (n == 0) return 0;
  size_t bytes = 0;
  WBFMM_CHECK(rocprim::radix_sort_keys(nullptr, bytes, list, out, n, 0, 3,
                                       stream()));
  void *tmp = nullptr;
  unsigned *dcnt = nullptr;
  if (dmalloc(&tmp, bytes) || dmalloc((void **)&dcnt, 8 * sizeof(unsigned)))
    return WBFMM_HIP_ERR_NOMEM;
  WBFMM_CHECK(rocprim::radix_sort_keys(tmp, bytes, list, out, n, 0, 3,
                                       stream()));
  WBFMM_CHECK(hipMemsetAsync(dcnt, 0, 8 * sizeof(unsigned), stream()));
  WBFMM_LAUNCH("tree_class_count", k_class_count,
               dim3((n + 255) / 256 < 512 ? (n + 255) / 256 : 512), dim3(256), 0, list, n,
               dcnt);
  WBFMM_CHECK(hipMemcpyAsync(ccount, dcnt, 8 * sizeof(unsigned),
                             hipMemcpyDeviceToHost, stream()));
  dfree(tmp);
  dfree(dcnt);
  WBFMM_CHECK(hipStreamSynchronize(stream()));
  unsigned s = 0;
  for (int c = 0; c < 8; c++) {
    cstart[c] = s;
    s += ccount[c];
  }
  return 0;
}

int wbfmm_hip_lap_shift_mfma(int mode, const uint32_t *parents, unsigned npar,
                             const uint32_t *cstart, const uint32_t *cend,
                             const uint32_t *cflag, int N, const double *in, double *out,
                             const void *plan_dev, const double *A1, const double *A3,
                             const double *A2, const double *CS,
                             const wbfmm_m2l_plan_t *plan_host) {
  return shift_entry<double>(mode, parents, npar, cstart, cend, cflag, N, in, out, plan_dev,
                             A1, A3, A2, CS, plan_host);
}

int wbfmm_hip_lap_shift_mfma_f32(int mode, const uint32_t *parents, unsigned npar,
                                 const uint32_t *cstart, const uint32_t *cend,
                                 const uint32_t *cflag, int N, const float *in, float *out,
                                 const void *plan_dev, const float *A1, const float *A3,
                                 const float *A2, const float *CS,
                                 const wbfmm_m2l_plan_t *plan_host) {
  return shift_entry<float>(mode, parents, npar, cstart, cend, cflag, N, in, out, plan_dev,
                            A1, A3, A2, CS, plan_host);
}

int wbfmm_hip_mfma_layout_probe(double *host_out) {
  double *d = nullptr;
  if (dmalloc((void **)&d, 256 * sizeof(double))) return WBFMM_HIP_ERR_NOMEM;
  WBFMM_LAUNCH("mfma_probe", k_mfma_layout_probe, dim3(1), dim3(64), 0, d);
  WBFMM_CHECK(hipMemcpyAsync(host_out, d, 256 * sizeof(double),
                             hipMemcpyDeviceToHost, stream()));
  dfree(d);
  WBFMM_CHECK(hipStreamSynchronize(stream()));
  return 0;
}

int wbfmm_hip_mfma_layout_probe_f32(float *host_out) {
  float *d = nullptr;
  if (dmalloc((void **)&d, 256 * sizeof(float))) return WBFMM_HIP_ERR_NOMEM;
  WBFMM_LAUNCH("mfma_probe_f32", k_mfma_layout_probe_f32, dim3(1), dim3(64), 0, d);
  WBFMM_CHECK(hipMemcpyAsync(host_out, d, 256 * sizeof(float),
                             hipMemcpyDeviceToHost, stream()));
  dfree(d);
  WBFMM_CHECK(hipStreamSynchronize(stream()));
  return 0;
}

int wbfmm_hip_lap_m2l_mfma(const uint32_t *cls, const unsigned *cstart,
                           const unsigned *ccount, const uint32_t *sstart,
                           const uint32_t *send, unsigned level, double D,
                           int N, const double *mps, double *mpr,
                           const void *plan_dev, const double *A1,
                           const double *A3, const double *A2,
                           const double *CS,
                           const wbfmm_m2l_plan_t *plan_host) {
  (void)D;
  return m2l_entry<double>(cls, cstart, ccount, sstart, send, level, N, mps, mpr, plan_dev, A1,
                           A3, A2, CS, plan_host);
}

int wbfmm_hip_lap_m2l_mfma_noalloc(const uint32_t *cls, const unsigned *cstart,
                                   const unsigned *ccount, const uint32_t *sstart,
                                   const uint32_t *send, unsigned level, double D,
                                   int N, const double *mps, double *mpr,
                                   const void *plan_dev, const double *A1,
                                   const double *A3, const double *A2,
                                   const double *CS,
                                   const wbfmm_m2l_plan_t *plan_host) {
  (void)D;
  unsigned char *ext = nullptr;
  const int rc = scratch_address(&ext);
  if (rc) return rc;
  return m2l_entry<double>(cls, cstart, ccount, sstart, send, level, N, mps, mpr, plan_dev, A1,
                           A3, A2, CS, plan_host, ext);
}

int wbfmm_hip_lap_m2l_mfma_f32(const uint32_t *cls, const unsigned *cstart,
                               const unsigned *ccount, const uint32_t *sstart,
                               const uint32_t *send, unsigned level, int N,
                               const float *mps, float *mpr, const void *plan_dev,
                               const float *A1, const float *A3, const float *A2,
                               const float *CS, const wbfmm_m2l_plan_t *plan_host) {
  return m2l_entry<float>(cls, cstart, ccount, sstart, send, level, N, mps, mpr, plan_dev, A1,
                          A3, A2, CS, plan_host);
}

}  // extern "C"
