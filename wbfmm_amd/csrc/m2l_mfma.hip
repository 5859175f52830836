// m2l_mfma.hip -- Laplace M2L as FP64 matrix-core products (gfx950).
//
// The reference translates one source box at a time: rotate the multipole
// into the frame of the offset, coaxial S|R shift, rotate back and add into
// the target's local expansion (src/pass-laplace.c:35-105).  Every box pair
// with the same offset shares the two rotations and every pair with the
// same distance shares the shift, so here a wavefront takes 16 target
// boxes with the same Morton child digit (hence the same 189 list-4
// offsets), and for each offset applies the three factors to the 16
// source expansions at once with v_mfma_f64_16x16x4f64:
//
//   stage 1  Y = R_o X        rows n-major, A from the offset's table
//   stage 2  Z = S Y          rows m-major, unit-distance S|R, inputs
//                             scaled by t^-nu and outputs by t^-(n+1)
//   stage 3  L += R_o^-1 Z    rows n-major, accumulated in registers
//
// X, Y and Z pass through one LDS buffer of (CP+4) x 16 doubles; the local
// expansions of the 16 targets stay in registers (NT double4 per lane)
// for all 189 offsets and are added to mpr once at the end.
//
// MFMA operand layout (16x16x4 f64, checked by wbfmm_hip_mfma_layout_probe):
// lane l supplies A[l%16][l/16] and B[l/16][l%16]; the result holds
// D[l/16 + 4q][l%16], q = 0..3.

#include <cstdio>
#include <cstring>

#include <hip/hip_runtime.h>
#include <rocprim/rocprim.hpp>

#include "hip_common.h"

namespace wbfmm {

typedef double d4 __attribute__((ext_vector_type(4)));

struct ClassArgs {
  unsigned start[8], count[8];
  unsigned chunk_prefix[9];
};

// Compile-time copy of the tiling built on the host by
// wbfmm_hip_lap_m2l_plan (the A tables are stored in this order); with it
// every k loop has a constant trip count and unrolls, so the A loads of a
// whole stage can be issued ahead of their products.
template <int N>
struct Tiling {
  static constexpr int C = (N + 1) * (N + 1);
  static constexpr int NT = (C + 15) / 16;
  int kb1[NT], nk1[NT], off1[NT], kb2[NT], nk2[NT], off2[NT];
  int T13, T2;
};

constexpr int tl_deg(int r) {
  int n = 0;
  while ((n + 1) * (n + 1) <= r) n++;
  return n;
}

// bounds of the m-major block holding m-major row r (order N)
constexpr int tl_mlo(int N, int r) {
  int start = 0, len = N + 1;
  if (r < start + len) return start;
  start += len;
  for (int m = 1; m <= N; m++)
    for (int part = 0; part < 2; part++) {
      len = N + 1 - m;
      if (r < start + len) return start;
      start += len;
    }
  return start;
}

constexpr int tl_mhi(int N, int r) {
  int start = 0, len = N + 1;
  if (r < start + len) return start + len;
  start += len;
  for (int m = 1; m <= N; m++)
    for (int part = 0; part < 2; part++) {
      len = N + 1 - m;
      if (r < start + len) return start + len;
      start += len;
    }
  return start;
}

template <int N>
constexpr Tiling<N> make_tiling() {
  Tiling<N> t{};
  constexpr int C = Tiling<N>::C, NT = Tiling<N>::NT;
  int a1 = 0, a2 = 0;
  for (int i = 0; i < NT; i++) {
    int r0 = 16 * i, r1 = (16 * i + 15 < C - 1) ? 16 * i + 15 : C - 1;
    int lo = tl_deg(r0) * tl_deg(r0), hi = (tl_deg(r1) + 1) * (tl_deg(r1) + 1);
    t.kb1[i] = lo;
    t.nk1[i] = (hi - lo + 3) / 4;
    t.off1[i] = a1;
    a1 += t.nk1[i];
    lo = tl_mlo(N, r0);
    hi = tl_mhi(N, r1);
    t.kb2[i] = lo;
    t.nk2[i] = (hi - lo + 3) / 4;
    t.off2[i] = a2;
    a2 += t.nk2[i];
  }
  t.T13 = a1;
  t.T2 = a2;
  return t;
}

template <int N>
__global__ __launch_bounds__(64) void k_lap_m2l_mfma(
    const uint32_t *__restrict__ cls, ClassArgs ca,
    const uint32_t *__restrict__ sstart, const uint32_t *__restrict__ send,
    int level, double w, const double *__restrict__ mps,
    double *__restrict__ mpr, const wbfmm_m2l_plan_t *__restrict__ P,
    const double *__restrict__ A1, const double *__restrict__ A3,
    const double *__restrict__ A2) {
  constexpr Tiling<N> TP = make_tiling<N>();
  constexpr int C = Tiling<N>::C;
  constexpr int NT = Tiling<N>::NT;
  constexpr int ROWS = 16 * NT + 4;
  __shared__ double buf[ROWS * 16];
  __shared__ double ipw[2 * N + 4];
  __shared__ int s_mpos[C], s_npos[C], s_degn[C], s_degm[C];
  const int lane = threadIdx.x, col = lane & 15, grp = lane >> 4;

  for (int r = lane; r < C; r += 64) {
    s_mpos[r] = P->mpos[r];
    s_npos[r] = P->npos[r];
    s_degn[r] = P->degn[r];
    s_degm[r] = P->degm[r];
  }

  int b = blockIdx.x, c = 0;
  while (c < 7 && (unsigned)b >= ca.chunk_prefix[c + 1]) c++;
  const unsigned k16 = (unsigned)(b - ca.chunk_prefix[c]) * 16 + col;
  const bool tvalid = k16 < ca.count[c];
  const uint32_t tb = tvalid ? cls[ca.start[c] + k16] : 0u;
  const int ti = compact3(tb), tj = compact3(tb >> 1), tk = compact3(tb >> 2);
  const int px = c & 1, py = (c >> 1) & 1, pz = (c >> 2) & 1;
  const int nbox = 1 << level;

  d4 acc[NT];
#pragma unroll
  for (int i = 0; i < NT; i++) acc[i] = d4{0.0, 0.0, 0.0, 0.0};

  for (int dx = -2 - px; dx <= 3 - px; dx++)
    for (int dy = -2 - py; dy <= 3 - py; dy++)
      for (int dz = -2 - pz; dz <= 3 - pz; dz++) {
        if (dx >= -1 && dx <= 1 && dy >= -1 && dy <= 1 && dz >= -1 && dz <= 1)
          continue;
        const int si = ti + dx, sj = tj + dy, sk = tk + dz;
        bool v = tvalid && si >= 0 && si < nbox && sj >= 0 && sj < nbox &&
                 sk >= 0 && sk < nbox;
        uint64_t sb = 0;
        if (v) {
          sb = spread3(si) | (spread3(sj) << 1) | (spread3(sk) << 2);
          v = send[sb] > sstart[sb];
        }
        if (!__any(v)) continue;
        const int o = (dx + 3) * 49 + (dy + 3) * 7 + dz + 3;
        const double lt = log(P->r[o] * w);
        __syncthreads();
        if (lane < 2 * N + 4) ipw[lane] = exp(-(double)lane * lt);
#pragma unroll
        for (int row = grp; row < ROWS; row += 4)
          buf[row * 16 + col] = (v && row < C) ? mps[coef_at(sb, row, C)] : 0.0;
        __syncthreads();

        // stage 1: rotate into the offset's frame
        d4 y[NT];
        const double *a1 = A1 + (size_t)o * TP.T13 * 64 + lane;
#pragma unroll
        for (int i = 0; i < NT; i++) {
          d4 r4 = d4{0.0, 0.0, 0.0, 0.0};
#pragma unroll
          for (int j = 0; j < TP.nk1[i]; j++)
            r4 = __builtin_amdgcn_mfma_f64_16x16x4f64(
                a1[(TP.off1[i] + j) * 64],
                buf[(TP.kb1[i] + 4 * j + grp) * 16 + col], r4, 0, 0, 0);
          y[i] = r4;
        }
        __syncthreads();
#pragma unroll
        for (int i = 0; i < NT; i++)
#pragma unroll
          for (int q = 0; q < 4; q++) {
            const int r = 16 * i + grp + 4 * q;
            if (r < C) buf[s_mpos[r] * 16 + col] = y[i][q] * ipw[s_degn[r]];
          }
        __syncthreads();

        // stage 2: coaxial S|R shift, m-major
        const double *a2 = A2 + lane;
#pragma unroll
        for (int i = 0; i < NT; i++) {
          d4 r4 = d4{0.0, 0.0, 0.0, 0.0};
#pragma unroll
          for (int j = 0; j < TP.nk2[i]; j++)
            r4 = __builtin_amdgcn_mfma_f64_16x16x4f64(
                a2[(TP.off2[i] + j) * 64],
                buf[(TP.kb2[i] + 4 * j + grp) * 16 + col], r4, 0, 0, 0);
          y[i] = r4;
        }
        __syncthreads();
#pragma unroll
        for (int i = 0; i < NT; i++)
#pragma unroll
          for (int q = 0; q < 4; q++) {
            const int r = 16 * i + grp + 4 * q;
            if (r < C) buf[s_npos[r] * 16 + col] = y[i][q] * ipw[s_degm[r] + 1];
          }
        __syncthreads();

        // stage 3: rotate back and accumulate
        const double *a3 = A3 + (size_t)o * TP.T13 * 64 + lane;
#pragma unroll
        for (int i = 0; i < NT; i++) {
#pragma unroll
          for (int j = 0; j < TP.nk1[i]; j++)
            acc[i] = __builtin_amdgcn_mfma_f64_16x16x4f64(
                a3[(TP.off1[i] + j) * 64],
                buf[(TP.kb1[i] + 4 * j + grp) * 16 + col], acc[i], 0, 0, 0);
        }
      }

  if (!tvalid) return;
#pragma unroll
  for (int i = 0; i < NT; i++)
#pragma unroll
    for (int q = 0; q < 4; q++) {
      const int r = 16 * i + grp + 4 * q;
      if (r < C) mpr[coef_at(tb, r, C)] += acc[i][q];
    }
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

__global__ void k_class_count(const uint32_t *list, unsigned n,
                              unsigned *counts) {
  unsigned i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) atomicAdd(&counts[list[i] & 7], 1u);
}

template <int N>
static bool tiling_matches(const wbfmm_m2l_plan_t *hp) {
  constexpr Tiling<N> t = make_tiling<N>();
  if (hp->N != N || hp->T13 != t.T13 || hp->T2 != t.T2) return false;
  for (int i = 0; i < Tiling<N>::NT; i++)
    if (hp->kb1[i] != t.kb1[i] || hp->nk1[i] != t.nk1[i] || hp->off1[i] != t.off1[i] ||
        hp->kb2[i] != t.kb2[i] || hp->nk2[i] != t.nk2[i] || hp->off2[i] != t.off2[i])
      return false;
  return true;
}

template <int N>
static int launch_m2l(const uint32_t *cls, const ClassArgs &ca, unsigned nchunks,
                      const uint32_t *sstart, const uint32_t *send, int level,
                      double w, const double *mps, double *mpr,
                      const void *plan, const double *A1, const double *A3,
                      const double *A2, const wbfmm_m2l_plan_t *hplan) {
  if (!tiling_matches<N>(hplan)) {
    fprintf(stderr, "wbfmm (HIP): M2L plan for order %d does not match the kernel's tiling\n", N);
    return WBFMM_HIP_ERR_ARG;
  }
  WBFMM_LAUNCH("lap_m2l", k_lap_m2l_mfma<N>, dim3(nchunks), dim3(64), 0, cls,
               ca, sstart, send, level, w, mps, mpr,
               (const wbfmm_m2l_plan_t *)plan, A1, A3, A2);
  return 0;
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
  WBFMM_CHECK(hipMalloc(&tmp, bytes ? bytes : 16));
  WBFMM_CHECK(hipMalloc(&dcnt, 8 * sizeof(unsigned)));
  WBFMM_CHECK(rocprim::radix_sort_keys(tmp, bytes, list, out, n, 0, 3,
                                       stream()));
  WBFMM_CHECK(hipMemsetAsync(dcnt, 0, 8 * sizeof(unsigned), stream()));
  WBFMM_LAUNCH("tree_class_count", k_class_count, dim3((n + 255) / 256),
               dim3(256), 0, list, n, dcnt);
  WBFMM_CHECK(hipMemcpyAsync(ccount, dcnt, 8 * sizeof(unsigned),
                             hipMemcpyDeviceToHost, stream()));
  WBFMM_CHECK(hipStreamSynchronize(stream()));
  (void)hipFree(tmp);
  (void)hipFree(dcnt);
  unsigned s = 0;
  for (int c = 0; c < 8; c++) {
    cstart[c] = s;
    s += ccount[c];
  }
  return 0;
}

int wbfmm_hip_mfma_layout_probe(double *host_out) {
  double *d = nullptr;
  WBFMM_CHECK(hipMalloc(&d, 256 * sizeof(double)));
  WBFMM_LAUNCH("mfma_probe", k_mfma_layout_probe, dim3(1), dim3(64), 0, d);
  WBFMM_CHECK(hipMemcpyAsync(host_out, d, 256 * sizeof(double),
                             hipMemcpyDeviceToHost, stream()));
  WBFMM_CHECK(hipStreamSynchronize(stream()));
  (void)hipFree(d);
  return 0;
}

int wbfmm_hip_lap_m2l_mfma(const uint32_t *cls, const unsigned *cstart,
                           const unsigned *ccount, const uint32_t *sstart,
                           const uint32_t *send, unsigned level, double D,
                           int N, const double *mps, double *mpr,
                           const void *plan_dev, const double *A1,
                           const double *A3, const double *A2,
                           const wbfmm_m2l_plan_t *plan_host) {
  ClassArgs ca;
  unsigned nch = 0;
  for (int c = 0; c < 8; c++) {
    ca.start[c] = cstart[c];
    ca.count[c] = ccount[c];
    ca.chunk_prefix[c] = nch;
    nch += (ccount[c] + 15) / 16;
  }
  ca.chunk_prefix[8] = nch;
  if (nch == 0) return 0;
  double w = D / (double)(1u << level);
  switch (N) {
#define CASE(n) \
  case n:       \
    return launch_m2l<n>(cls, ca, nch, sstart, send, (int)level, w, mps, mpr, plan_dev, A1, A3, A2, plan_host);
    CASE(2) CASE(3) CASE(4) CASE(5) CASE(6) CASE(7) CASE(8) CASE(9) CASE(10)
    CASE(11) CASE(12) CASE(13) CASE(14) CASE(15) CASE(16)
#undef CASE
    default:
      return WBFMM_HIP_ERR_ORDER;
  }
}

}  // extern "C"
