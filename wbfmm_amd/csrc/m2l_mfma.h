// m2l_mfma.h -- the bin-packed MFMA plan shared by the Laplace and
// Helmholtz matrix-core translation kernels (m2l_mfma.hip,
// hel_m2l_mfma.hip): operand traits, the compile-time copy of the host bin
// packing, the per-wavefront stage product and the split reduction.

#ifndef WBFMM_M2L_MFMA_H
#define WBFMM_M2L_MFMA_H

#include <hip/hip_runtime.h>

#include "hip_common.h"

namespace wbfmm {

typedef double d4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));

// The kernels run in FP64 (v_mfma_f64_16x16x4f64) or FP32
// (v_mfma_f32_16x16x4f32).  Both take A[l%16][l/16] and B[l/16][l%16] from
// lane l; the four results of a lane are rows drow(l/16, q) of column l%16
// (checked by wbfmm_hip_mfma_layout_probe and _probe_f32).
template <typename T> struct MF;
template <> struct MF<double> {
  typedef d4 v4;
  typedef double2 t2;
  static __device__ __forceinline__ d4 mma(double a, double b, d4 c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  static constexpr __host__ __device__ int drow(int g, int q) { return g + 4 * q; }
};
template <> struct MF<float> {
  typedef f4 v4;
  typedef float2 t2;
  static __device__ __forceinline__ f4 mma(float a, float b, f4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  static constexpr __host__ __device__ int drow(int g, int q) { return 4 * g + q; }
};
template <typename T>
__device__ __forceinline__ typename MF<T>::t2 mk2(T a, T b) {
  typename MF<T>::t2 v;
  v.x = a;
  v.y = b;
  return v;
}

struct ClassArgs {
  unsigned start[8], count[8];
  unsigned chunk_prefix[9];
  // chunks per class and XCD when the blocks interleave the classes (see
  // block_chunk); 0: chunks in class order
  unsigned qc;
  // the chunked Laplace M2L's order of the offsets: 0 cube index, 1 offset
  // class (kOffOrder)
  unsigned ord = 0;
};

// The chunk (class c, index j within the class) of block bid, or false
// for a block with none.  Blocks go round-robin over the 8 XCDs (block
// bid on XCD bid % 8).  qc == 0: chunks in class order, consecutive chunks
// of one range on each XCD.  qc > 0: XCD x takes chunks x qc .. x qc + qc
// - 1 of every class, the 8 classes interleaved, so the blocks an XCD
// runs at one time cover one compact region of the level and share its
// source expansions in that XCD's L2 (grid: 64 qc blocks).
__device__ __forceinline__ bool block_chunk(unsigned bid, const ClassArgs &ca, int &c,
                                            unsigned &j) {
  if (ca.qc == 0) {
    const unsigned nch = ca.chunk_prefix[8], q8 = nch / 8;
    const unsigned chunk = bid < 8 * q8 ? (bid & 7) * q8 + (bid >> 3) : bid;
    c = 0;
    while (c < 7 && chunk >= ca.chunk_prefix[c + 1]) c++;
    j = chunk - ca.chunk_prefix[c];
    return chunk < nch;
  }
  const unsigned x = bid & 7, local = bid >> 3, jj = local >> 3;
  c = (int)(local & 7);
  j = x * ca.qc + jj;
  return jj < ca.qc && j < ca.chunk_prefix[c + 1] - ca.chunk_prefix[c];
}

// Compile-time copy of the bin packing of wbfmm_tables.c (pack_bins over
// the cos/sin blocks, sizes n+1 and n): number of bins and the k steps of
// each bin, so that every MFMA loop has a constant trip count.
struct Bins {
  int nb, kt;
  int nk[WBFMM_M2L_MAXB], koff[WBFMM_M2L_MAXB], kbase[WBFMM_M2L_MAXB];
};

// blocks of more than 16 rows take consecutive bins of their own, whose
// row tiles all read the block's full column range from its first row
constexpr Bins make_bins(int N) {
  Bins B{};
  int sz[2 * WBFMM_M2L_MAXN + 2] = {}, order[2 * WBFMM_M2L_MAXN + 2] = {};
  int fill[WBFMM_M2L_MAXB] = {}, kb[WBFMM_M2L_MAXB] = {}, nk[WBFMM_M2L_MAXB] = {};
  const int nblk = 2 * N + 1;
  for (int i = 0; i < nblk; i++) {
    sz[i] = i <= N ? i + 1 : i - N;
    order[i] = i;
  }
  for (int i = 1; i < nblk; i++) {
    int v = order[i], j = i;
    for (; j > 0 && sz[order[j - 1]] < sz[v]; j--) order[j] = order[j - 1];
    order[j] = v;
  }
  int nb = 0;
  for (int i = 0; i < nblk; i++) {
    int k = order[i], b = 0;
    if (sz[k] > 16) {
      const int nt = (sz[k] + 15) / 16;
      for (int t = 0; t < nt; t++) {
        fill[nb + t] = 16;
        kb[nb + t] = 16 * nb;
        nk[nb + t] = (sz[k] + 3) / 4;
      }
      nb += nt;
      continue;
    }
    for (; b < nb; b++)
      if (fill[b] + sz[k] <= 16) break;
    if (b == nb) {
      fill[nb] = 0;
      kb[nb] = -1;
      nb++;
    }
    fill[b] += sz[k];
  }
  B.nb = nb;
  for (int b = 0; b < nb; b++) {
    if (kb[b] < 0) {
      kb[b] = 16 * b;
      nk[b] = (fill[b] + 3) / 4;
    }
    B.nk[b] = nk[b];
    B.kbase[b] = kb[b];
    B.koff[b] = B.kt;
    B.kt += B.nk[b];
  }
  return B;
}

template <int N>
struct M2LShape {
  static constexpr Bins BN = make_bins(N);
  static constexpr int NB = BN.nb;
  static constexpr int KT = BN.kt;
  static constexpr int C = (N + 1) * (N + 1);
  static constexpr int NU = (N + 1) * (N + 2) / 2;
  static constexpr int S = (16 * NB > 2 * NU) ? 16 * NB : 2 * NU;  // = plan S
  static constexpr int NAU = (NU + 3) / 4;         // units per lane
  // register prefetch of the next offset's sources while it fits
  static constexpr bool PREFETCH = N <= 12;
};

// one stage: y[b] = tile(b) * R[bin b's k slots].  The A operands (lane
// order) run through a ring of AD registers: step t is requested AD steps
// before its product, so a product waits on a load issued long before.
// `a` is wave-uniform (scalar base) and the lane index is added per load:
// the loads use a scalar base plus a 32-bit lane offset.
template <typename T, int NB, int KT>
__device__ __forceinline__ void stage(const Bins &BN, const T *__restrict__ a,
                                      const T *R, int lane, int grp, int col,
                                      typename MF<T>::v4 *y) {
  constexpr int AD = KT < 16 ? KT : 16;
  T av[AD];
#pragma unroll
  for (int t = 0; t < AD; t++) av[t] = a[t * 64 + lane];
#pragma unroll
  for (int b = 0; b < NB; b++) {
    typename MF<T>::v4 r4 = {T(0), T(0), T(0), T(0)};
#pragma unroll
    for (int j = 0; j < BN.nk[b]; j++) {
      const int t = BN.koff[b] + j;
      r4 = MF<T>::mma(av[t % AD], R[(BN.kbase[b] + 4 * j + grp) * 16 + col], r4);
      if (t + AD < KT) av[t % AD] = a[(t + AD) * 64 + lane];
    }
    y[b] = r4;
  }
}

// k-steps of the bins b = WI (mod W), in bin order
struct WaveSeq {
  int n, nb;
  int t[8 * WBFMM_M2L_MAXB];
  int start[WBFMM_M2L_MAXB];
};

constexpr WaveSeq make_seq(const Bins &B, int W, int WI) {
  WaveSeq s{};
  for (int b = WI; b < B.nb; b += W) {
    s.start[s.nb++] = s.n;
    for (int j = 0; j < B.nk[b]; j++) s.t[s.n++] = B.koff[b] + j;
  }
  return s;
}

template <typename T, int N, int W, int WI>
__device__ __forceinline__ void stage_w(const T *__restrict__ a, const T *R, int lane, int grp,
                                        int col, typename MF<T>::v4 *y) {
  constexpr Bins BN = M2LShape<N>::BN;
  constexpr WaveSeq WS = make_seq(BN, W, WI);
  // a wavefront may have no bins at low orders (AD then only sizes av)
  constexpr int AD = WS.n < 16 ? (WS.n > 0 ? WS.n : 1) : 16;
  T av[AD];
#pragma unroll
  for (int s = 0; s < AD && s < WS.n; s++) av[s] = a[WS.t[s] * 64 + lane];
#pragma unroll
  for (int lb = 0; lb < WS.nb; lb++) {
    const int b = WI + W * lb;
    typename MF<T>::v4 r4 = {T(0), T(0), T(0), T(0)};
#pragma unroll
    for (int j = 0; j < BN.nk[b]; j++) {
      const int s = WS.start[lb] + j;
      r4 = MF<T>::mma(av[s % AD], R[(BN.kbase[b] + 4 * j + grp) * 16 + col], r4);
      if (s + AD < WS.n) av[s % AD] = a[WS.t[s + AD] * 64 + lane];
    }
    y[lb] = r4;
  }
}

// one stage with its results stored at once into the other LDS buffer
// (the double-buffered form: the stage reads Rin and writes Rout, so no
// barrier is needed between its reads and its writes)
template <typename T, int N, int W, int WI>
__device__ __forceinline__ void stage_wo(const T *__restrict__ a, const T *Rin, T *Rout,
                                         const int4 *tm, int lane, int grp, int col) {
  constexpr Bins BN = M2LShape<N>::BN;
  constexpr WaveSeq WS = make_seq(BN, W, WI);
  constexpr int AD = WS.n < 16 ? (WS.n > 0 ? WS.n : 1) : 16;
  T av[AD];
#pragma unroll
  for (int s = 0; s < AD && s < WS.n; s++) av[s] = a[WS.t[s] * 64 + lane];
#pragma unroll
  for (int lb = 0; lb < WS.nb; lb++) {
    const int b = WI + W * lb;
    typename MF<T>::v4 r4 = {T(0), T(0), T(0), T(0)};
#pragma unroll
    for (int j = 0; j < BN.nk[b]; j++) {
      const int s = WS.start[lb] + j;
      r4 = MF<T>::mma(av[s % AD], Rin[(BN.kbase[b] + 4 * j + grp) * 16 + col], r4);
      if (s + AD < WS.n) av[s % AD] = a[WS.t[s + AD] * 64 + lane];
    }
    const int4 d = tm[4 * b + grp];
    Rout[d.x * 16 + col] = r4[0];
    Rout[d.y * 16 + col] = r4[1];
    Rout[d.z * 16 + col] = r4[2];
    Rout[d.w * 16 + col] = r4[3];
  }
}

// Operands held across stages: where a wavefront's k-steps of a stage fit
// in 16 registers (pairs), av[s] holds the A operand of its step s.  A
// stage consumes av[s] and at once requests step s of the next stage's
// table into it, so every stage starts with its operands loaded, or in
// flight since the stage before, instead of waiting for them after the
// barrier that precedes it.
template <int N, int W, int WI>
struct StageOps {
  static constexpr WaveSeq WS = make_seq(M2LShape<N>::BN, W, WI);
  static constexpr int n = WS.n > 0 ? WS.n : 1;
  static constexpr bool fits = WS.n <= 16;
};

template <typename T, int N, int W, int WI>
__device__ __forceinline__ void ops_load(T *av, const T *__restrict__ a, int lane) {
  constexpr WaveSeq WS = make_seq(M2LShape<N>::BN, W, WI);
#pragma unroll
  for (int s = 0; s < WS.n; s++) av[s] = a[WS.t[s] * 64 + lane];
}

// stage_w with the operands in av; anext: the next stage's table
template <typename T, int N, int W, int WI>
__device__ __forceinline__ void stage_wp(T *av, const T *__restrict__ anext, const T *R,
                                         int lane, int grp, int col, typename MF<T>::v4 *y) {
  constexpr Bins BN = M2LShape<N>::BN;
  constexpr WaveSeq WS = make_seq(BN, W, WI);
#pragma unroll
  for (int lb = 0; lb < WS.nb; lb++) {
    const int b = WI + W * lb;
    typename MF<T>::v4 r4 = {T(0), T(0), T(0), T(0)};
#pragma unroll
    for (int j = 0; j < BN.nk[b]; j++) {
      const int s = WS.start[lb] + j;
      r4 = MF<T>::mma(av[s], R[(BN.kbase[b] + 4 * j + grp) * 16 + col], r4);
      av[s] = anext[WS.t[s] * 64 + lane];
    }
    y[lb] = r4;
  }
}

// stage_wo with the operands in av; anext: the next stage's table
template <typename T, int N, int W, int WI>
__device__ __forceinline__ void stage_wop(T *av, const T *__restrict__ anext, const T *Rin,
                                          T *Rout, const int4 *tm, int lane, int grp, int col) {
  constexpr Bins BN = M2LShape<N>::BN;
  constexpr WaveSeq WS = make_seq(BN, W, WI);
#pragma unroll
  for (int lb = 0; lb < WS.nb; lb++) {
    const int b = WI + W * lb;
    typename MF<T>::v4 r4 = {T(0), T(0), T(0), T(0)};
#pragma unroll
    for (int j = 0; j < BN.nk[b]; j++) {
      const int s = WS.start[lb] + j;
      r4 = MF<T>::mma(av[s], Rin[(BN.kbase[b] + 4 * j + grp) * 16 + col], r4);
      av[s] = anext[WS.t[s] * 64 + lane];
    }
    const int4 d = tm[4 * b + grp];
    Rout[d.x * 16 + col] = r4[0];
    Rout[d.y * 16 + col] = r4[1];
    Rout[d.z * 16 + col] = r4[2];
    Rout[d.w * 16 + col] = r4[3];
  }
}

// the exchange buffer is doubled where the two copies take at most 48 KB,
// which leaves room for three blocks per CU (FP64 up to order 12, FP32 up
// to order 17)
template <typename T, int N>
struct M2LDouble {
  static constexpr bool value = 2 * (M2LShape<N>::S + 1) * 16 * sizeof(T) <= 48 * 1024;
};

// mpr[cls[i]] += sum over the splits (in split order) of the partials of
// target i
template <typename T>
__global__ void k_m2l_reduce(const uint32_t *__restrict__ cls, unsigned ntot, int nsplit,
                             int C, const T *__restrict__ part, T *__restrict__ mpr) {
  const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (size_t)ntot * C) return;
  const size_t i = e / C, c = e % C;
  T s = part[i * C + c];
  for (int sp = 1; sp < nsplit; sp++) s += part[((size_t)sp * ntot + i) * C + c];
  mpr[(size_t)cls[i] * C + c] += s;
}

template <int N>
inline bool plan_matches(const wbfmm_m2l_plan_t *hp) {
  constexpr Bins B = make_bins(N);
  if (hp->N != N || hp->NB != B.nb || hp->KT != B.kt || hp->S != M2LShape<N>::S) return false;
  for (int b = 0; b < B.nb; b++)
    if (hp->nk[b] != B.nk[b] || hp->koff[b] != B.koff[b] || hp->kbase[b] != B.kbase[b])
      return false;
  return true;
}

// a zero expansion of 2 (WBFMM_M2L_MAXN+1)^2 doubles read by the kernels
// in place of absent boxes; made on first use, read-only afterwards
int ensure_zero();
const double *zero_expansion();
// occupancy bitmap of a level (bit b: box b holds points), stream-ordered
// memory released by the caller with dfree
int occupancy_bits(const uint32_t *sstart, const uint32_t *send, unsigned level,
                   uint32_t **occ);

// Offset classes of list 4 for the packed sparse-level M2L kernels (Laplace
// and Helmholtz): offsets with the same dz and the same dx^2 + dy^2 share
// the polar angle and the distance, hence all three stage tables.  Per
// parity c = px | py << 1 | pz << 2 of the targets' child digit, the classes
// in the order of the key (dz + 3) 32 + dx^2 + dy^2, members in cube-index
// order (k = 36 (dz+2+pz) + 6 (dy+2+py) + dx+2+px).
constexpr int kMaxOffCls = 64;

struct OffClasses {
  int n;                        // classes (51 for every parity)
  uint8_t nm[kMaxOffCls];       // members per class (at most 8)
  uint8_t k[kMaxOffCls][8];     // members' cube indices, increasing
};

constexpr OffClasses make_off_classes(int c) {
  OffClasses O{};
  const int px = c & 1, py = (c >> 1) & 1, pz = (c >> 2) & 1;
  for (int key = 0; key < 7 * 32; key++) {
    int m = 0;
    for (int k = 0; k < 216; k++) {
      const int dx = k % 6 - 2 - px, dy = (k / 6) % 6 - 2 - py, dz = k / 36 - 2 - pz;
      const bool nr = dx >= -1 && dx <= 1 && dy >= -1 && dy <= 1 && dz >= -1 && dz <= 1;
      if (nr || (dz + 3) * 32 + dx * dx + dy * dy != key) continue;
      O.k[O.n][m++] = (uint8_t)k;
    }
    if (m) O.nm[O.n++] = (uint8_t)m;
  }
  return O;
}

static __constant__ OffClasses kOffCls[8] = {make_off_classes(0), make_off_classes(1),
                                             make_off_classes(2), make_off_classes(3),
                                             make_off_classes(4), make_off_classes(5),
                                             make_off_classes(6), make_off_classes(7)};
static_assert(make_off_classes(0).n <= kMaxOffCls && make_off_classes(7).n <= kMaxOffCls,
              "offset classes");

// the far offsets of each parity in class order (the members of
// kOffCls[c] one class after the other): 189 cube indices
struct OffOrder {
  uint8_t k[189];
};

constexpr OffOrder make_off_order(int c) {
  OffOrder O{};
  const OffClasses C = make_off_classes(c);
  int n = 0;
  for (int g = 0; g < C.n; g++)
    for (int i = 0; i < C.nm[g]; i++) O.k[n++] = C.k[g][i];
  return O;
}

static __constant__ OffOrder kOffOrder[8] = {make_off_order(0), make_off_order(1),
                                             make_off_order(2), make_off_order(3),
                                             make_off_order(4), make_off_order(5),
                                             make_off_order(6), make_off_order(7)};

}  // namespace wbfmm

#endif
