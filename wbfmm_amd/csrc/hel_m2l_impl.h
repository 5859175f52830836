// hel_m2l_impl.h -- the Helmholtz matrix-core translation kernels (M2L,
// M2M, L2L) and their launchers, as templates over the precision and the
// expansion order.  Each hel_m2l_o*.hip translation unit instantiates one
// range of orders (WBFMM_HEL_RANGES), so the orders compile in parallel;
// hel_m2l_mfma.hip holds the entry points, the design notes and the
// dispatch by order.

#ifndef WBFMM_HEL_M2L_IMPL_H
#define WBFMM_HEL_M2L_IMPL_H

#include <cstdio>
#include <cstdlib>

#include <hip/hip_runtime.h>

#include "hip_common.h"
#include "m2l_mfma.h"
#include "hel_m2l_ranges.h"

namespace wbfmm {

// coefficient types of a precision: complex pairs and MFMA result vectors
template <typename T>
using T2t = typename MF<T>::t2;
template <typename T>
using V4t = typename MF<T>::v4;

// stage 2: complex coaxial tiles ar + i ai.  The imaginary product reads
// the partner column, so that column 2t receives Ar Zr - Ai Zi and column
// 2t+1 receives Ar Zi + Ai Zr.
template <typename T, int N, int W, int WI>
__device__ __forceinline__ void stage_cx(const T *__restrict__ ar,
                                         const T *__restrict__ ai, const T *R,
                                         int lane, int grp, int col, V4t<T> *y) {
  constexpr Bins BN = M2LShape<N>::BN;
  constexpr WaveSeq WS = make_seq(BN, W, WI);
  constexpr int AD = WS.n < 8 ? (WS.n > 0 ? WS.n : 1) : 8;
  T avr[AD], avi[AD];
#pragma unroll
  for (int s = 0; s < AD && s < WS.n; s++) {
    avr[s] = ar[WS.t[s] * 64 + lane];
    avi[s] = ai[WS.t[s] * 64 + lane];
  }
  const T sg = (col & 1) ? T(1) : T(-1);
  const int pc = col ^ 1;
#pragma unroll
  for (int lb = 0; lb < WS.nb; lb++) {
    const int b = WI + W * lb;
    V4t<T> r4 = {T(0), T(0), T(0), T(0)}, i4 = {T(0), T(0), T(0), T(0)};
#pragma unroll
    for (int j = 0; j < BN.nk[b]; j++) {
      const int s = WS.start[lb] + j;
      const int row = (BN.kbase[b] + 4 * j + grp) * 16;
      r4 = MF<T>::mma(avr[s % AD], R[row + col], r4);
      i4 = MF<T>::mma(avi[s % AD], R[row + pc], i4);
      if (s + AD < WS.n) {
        avr[s % AD] = ar[WS.t[s + AD] * 64 + lane];
        avi[s % AD] = ai[WS.t[s + AD] * 64 + lane];
      }
    }
    y[lb] = r4 + sg * i4;
  }
}

// stage 2 of the T-buffered form: as stage_cx, results stored per bin
// into the other buffer
template <typename T, int N, int W, int WI>
__device__ __forceinline__ void stage_cx_wo(const T *__restrict__ ar,
                                            const T *__restrict__ ai, const T *Rin,
                                            T *Rout, const int4 *tm, int lane, int grp,
                                            int col) {
  constexpr Bins BN = M2LShape<N>::BN;
  constexpr WaveSeq WS = make_seq(BN, W, WI);
  constexpr int AD = WS.n < 8 ? (WS.n > 0 ? WS.n : 1) : 8;
  T avr[AD], avi[AD];
#pragma unroll
  for (int s = 0; s < AD && s < WS.n; s++) {
    avr[s] = ar[WS.t[s] * 64 + lane];
    avi[s] = ai[WS.t[s] * 64 + lane];
  }
  const T sg = (col & 1) ? T(1) : T(-1);
  const int pc = col ^ 1;
#pragma unroll
  for (int lb = 0; lb < WS.nb; lb++) {
    const int b = WI + W * lb;
    V4t<T> r4 = {T(0), T(0), T(0), T(0)}, i4 = {T(0), T(0), T(0), T(0)};
#pragma unroll
    for (int j = 0; j < BN.nk[b]; j++) {
      const int s = WS.start[lb] + j;
      const int row = (BN.kbase[b] + 4 * j + grp) * 16;
      r4 = MF<T>::mma(avr[s % AD], Rin[row + col], r4);
      i4 = MF<T>::mma(avi[s % AD], Rin[row + pc], i4);
      if (s + AD < WS.n) {
        avr[s % AD] = ar[WS.t[s + AD] * 64 + lane];
        avi[s % AD] = ai[WS.t[s + AD] * 64 + lane];
      }
    }
    const V4t<T> y = r4 + sg * i4;
    const int4 d = tm[4 * b + grp];
    Rout[d.x * 16 + col] = y[0];
    Rout[d.y * 16 + col] = y[1];
    Rout[d.z * 16 + col] = y[2];
    Rout[d.w * 16 + col] = y[3];
  }
}

template <typename T, int N, int W, int WI>
__device__ __forceinline__ void hel_m2l_loop(
    T *R, const int4 *tmap1, const int4 *tmap2, const int4 *tmap3, const int4 *uinf,
    T2t<T> (*cs)[N + 1], unsigned (*vm)[7], const unsigned *anym, int lane, int col, int grp,
    bool tvalid, uint32_t tb, int ti, int tj, int tk, int px, int py, int pz, int c,
    unsigned k8, int split, int nsplit, const ClassArgs &ca,
    const wbfmm_m2l_plan_t *__restrict__ P, const T *__restrict__ A,
    const T *__restrict__ A2, const T *__restrict__ CS,
    const T2t<T> *__restrict__ zero, const T2t<T> *__restrict__ mps,
    T2t<T> *__restrict__ mpr, T2t<T> *__restrict__ part, unsigned ntot) {
  using SH = M2LShape<N>;
  constexpr int C = SH::C, NU = SH::NU, KT = SH::KT;
  constexpr int NAW = (NU + 4 * W - 1) / (4 * W);  // units per lane
  constexpr WaveSeq WS = make_seq(SH::BN, W, WI);
  constexpr int NBW = WS.nb;
  constexpr bool DB = M2LDouble<T, N>::value;
  T *R1 = R + (SH::S + 1) * 16;  // the second buffer (DB only)
  const int ubase = grp + 4 * WI, tg = col >> 1;

  // the used offsets of this block's split as scalar bit words; the next
  // one after k by a scan of the words (216 when there is none)
  unsigned am[7];
#pragma unroll
  for (int w = 0; w < 7; w++) {
    unsigned v = __builtin_amdgcn_readfirstlane(anym[w]);
    if (nsplit > 1)
      for (int b = 0; b < 32; b++)
        if ((32 * w + b) % nsplit != split) v &= ~(1u << b);
    am[w] = v;
  }
  auto next_k = [&](int k) {
    int kn = k + 1;
#pragma unroll
    for (int w = 0; w < 7; w++) {
      if (kn < 32 * (w + 1)) {
        const unsigned bits = am[w] & (kn > 32 * w ? ~0u << (kn - 32 * w) : ~0u);
        if (bits) return 32 * w + (int)__builtin_ctz(bits);
        kn = 32 * (w + 1);
      }
    }
    return 216;
  };
  auto offset_of = [&](int k) {
    const int dx = k % 6 - 2 - px, dy = (k / 6) % 6 - 2 - py, dz = k / 36 - 2 - pz;
    return (dx + 3) * 49 + (dy + 3) * 7 + dz + 3;
  };
  // every wavefront writes the offset's factors (the same values)
  auto source_of = [&](int k, int p) {
    const int dx = k % 6 - 2 - px, dy = (k / 6) % 6 - 2 - py, dz = k / 36 - 2 - pz;
    if (lane <= N) {
      const T *t = CS + ((size_t)offset_of(k) * (N + 1) + lane) * 2;
      cs[p][lane] = mk2<T>(t[0], t[1]);
    }
    const bool v = (vm[tg][k >> 5] >> (k & 31)) & 1u;
    const uint32_t sb = morton_add(tb, dilate_offset(dx, 0, kMortonX),
                                   dilate_offset(dy, 1, kMortonY), dilate_offset(dz, 2, kMortonZ));
    return v ? mps + (uint64_t)sb * C : zero;
  };
  T pre_re[NAW], pre_im[NAW];
  // x_{+m} on the real lane, x_{-m} on the imaginary one
  auto load_sources = [&](const T2t<T> *src, const int4 *tu, int ub, int sp) {
#pragma unroll
    for (int i = 0; i < NAW; i++) {
      const int u = min(ub + 4 * W * i, NU - 1);
      const int4 inf = tu[u];
      const T2t<T> q = src[inf.x + (sp ? -inf.y : inf.y)];
      pre_re[i] = q.x;
      pre_im[i] = q.y;
    }
  };
  // x'_{+-m} = e^{+-i m chi} x_{+-m}; the partner lane holds the other
  // sign.  Lanes past the last unit (loaded as unit NU-1) exchange but
  // write nothing.
  auto deposit = [&](const int4 *tu, int ub, int cl, int p) {
    const int sp = cl & 1;
#pragma unroll
    for (int i = 0; i < NAW; i++) {
      const int u = min(ub + 4 * W * i, NU - 1);
      const int4 inf = tu[u];
      const T2t<T> e = cs[p][inf.y];
      const T es = sp ? -e.y : e.y;
      const T xr = e.x * pre_re[i] - es * pre_im[i];
      const T xi = es * pre_re[i] + e.x * pre_im[i];
      const T yr = __shfl_xor(xr, 1), yi = __shfl_xor(xi, 1);
      if (ub + 4 * W * i >= NU) continue;
      // real lane: x' = x'_{+m}, partner x'_{-m}; imaginary lane: the reverse
      R[inf.z * 16 + cl] = sp ? yi + xi : xr + yr;
      R[inf.w * 16 + cl] = sp ? yi - xi : xr - yr;
    }
  };
  auto write_out = [&](const int4 *tm, const V4t<T> *y, int gp, int cl) {
#pragma unroll
    for (int lb = 0; lb < NBW; lb++) {
      const int4 d = tm[4 * (WI + W * lb) + gp];
      R[d.x * 16 + cl] = y[lb][0];
      R[d.y * 16 + cl] = y[lb][1];
      R[d.z * 16 + cl] = y[lb][2];
      R[d.w * 16 + cl] = y[lb][3];
    }
  };

  T acc_re[NAW], acc_im[NAW];
#pragma unroll
  for (int i = 0; i < NAW; i++) acc_re[i] = acc_im[i] = T(0);

  // k is the same in every lane (it comes from the shared anym): held as a
  // scalar, and the next offset's table indices are read one offset ahead
  int k = __builtin_amdgcn_readfirstlane(next_k(-1)), p = 0;
  int th = 0, ri = 0;
  if (k < 216) {
    th = __builtin_amdgcn_readfirstlane(P->th[offset_of(k)]);
    ri = __builtin_amdgcn_readfirstlane(P->rix[offset_of(k)]);
    load_sources(source_of(k, 0), uinf, ubase, col & 1);
    deposit(uinf, ubase, col, 0);
  }
  __syncthreads();
  while (k < 216) {
    int z0;
    asm volatile("s_mov_b32 %0, 0" : "=s"(z0));
    const int4 *t1 = tmap1 + z0, *t2 = tmap2 + z0, *t3 = tmap3 + z0, *tu = uinf + z0;
    const int ln = lane + z0, cl = ln & 15, gp = ln >> 4, ub = gp + 4 * WI, sp = cl & 1;
    const int kn = __builtin_amdgcn_readfirstlane(next_k(k));
    int th_n = 0, ri_n = 0;
    if (kn < 216) {
      th_n = __builtin_amdgcn_readfirstlane(P->th[offset_of(kn)]);
      ri_n = __builtin_amdgcn_readfirstlane(P->rix[offset_of(kn)]);
    }

    // T-buffered where it fits: R -> R1 -> R -> R1, one barrier per
    // stage, and the next deposit goes into R while R1 is read below
    T *RZ = R;
    if constexpr (DB) {
      stage_wo<T, N, W, WI>(A + (size_t)th * KT * 64, R, R1, t1, ln, gp, cl);
      __syncthreads();
      if (kn < 216) load_sources(source_of(kn, p ^ 1), tu, ub, sp);
      stage_cx_wo<T, N, W, WI>(A2 + (size_t)(2 * ri) * KT * 64, A2 + (size_t)(2 * ri + 1) * KT * 64,
                            R1, R, t2, ln, gp, cl);
      __syncthreads();
      stage_wo<T, N, W, WI>(A + (size_t)th * KT * 64, R, R1, t3, ln, gp, cl);
      __syncthreads();
      RZ = R1;
    } else {
      V4t<T> y[NBW > 0 ? NBW : 1];
      stage_w<T, N, W, WI>(A + (size_t)th * KT * 64, R, ln, gp, cl, y);
      __syncthreads();
      write_out(t1, y, gp, cl);
      __syncthreads();
      if (kn < 216) load_sources(source_of(kn, p ^ 1), tu, ub, sp);

      stage_cx<T, N, W, WI>(A2 + (size_t)(2 * ri) * KT * 64, A2 + (size_t)(2 * ri + 1) * KT * 64,
                         R, ln, gp, cl, y);
      __syncthreads();
      write_out(t2, y, gp, cl);
      __syncthreads();

      stage_w<T, N, W, WI>(A + (size_t)th * KT * 64, R, ln, gp, cl, y);
      __syncthreads();
      write_out(t3, y, gp, cl);
      __syncthreads();
    }

    // y_{+-nu} = (P +- Q)/2 (y_0 = P/2), then L^{+-nu} += e^{-+i nu chi} y
    {
      const int c0 = cl & ~1;
#pragma unroll
      for (int i = 0; i < NAW; i++) {
        const int u = min(ub + 4 * W * i, NU - 1);
        const int nu = tu[u].y;
        const T2t<T> e = cs[p][nu];
        const T es = sp ? -e.y : e.y;
        const T pr = RZ[u * 16 + c0], pi = RZ[u * 16 + c0 + 1];
        const T qr = nu ? RZ[(NU + u) * 16 + c0] : T(0);
        const T qi = nu ? RZ[(NU + u) * 16 + c0 + 1] : T(0);
        const T yr = T(0.5) * (sp ? pr - qr : pr + qr);
        const T yi = T(0.5) * (sp ? pi - qi : pi + qi);
        acc_re[i] += e.x * yr + es * yi;
        acc_im[i] += e.x * yi - es * yr;
      }
    }
    if constexpr (!DB) __syncthreads();  // every read of R done before the next deposit

    if (kn < 216) deposit(tu, ub, cl, p ^ 1);
    __syncthreads();
    p ^= 1;
    k = kn;
    th = th_n;
    ri = ri_n;
  }

  if (!tvalid) return;
  const int sp = col & 1;
  T2t<T> *dst = nsplit > 1 ? part + ((size_t)split * ntot + ca.start[c] + k8) * C
                            : mpr + (size_t)tb * C;
#pragma unroll
  for (int i = 0; i < NAW; i++) {
    const int u = ubase + 4 * W * i;
    if (u >= NU) continue;
    const int4 inf = uinf[u];
    if (sp && inf.y == 0) continue;  // L^0 is the real lane's
    const int r = inf.x + (sp ? -inf.y : inf.y);
    if (nsplit > 1) {
      dst[r] = mk2<T>(acc_re[i], acc_im[i]);
    } else {
      const T2t<T> o = dst[r];
      dst[r] = mk2<T>(o.x + acc_re[i], o.y + acc_im[i]);
    }
  }
}

// FP32 from order 16 (eight wavefronts): held to 128 registers so that two
// blocks share a CU (the LDS buffer leaves room for two; at 159 registers
// only one fits)
template <typename T, int N, int W>
constexpr int hel_m2l_waves_per_eu() {
  return sizeof(T) == 4 && W == 8 && N >= 16 ? 4 : 1;
}

template <typename T, int N, int W>
__global__ __launch_bounds__(64 * W) __attribute__((amdgpu_waves_per_eu(
    hel_m2l_waves_per_eu<T, N, W>()))) void k_hel_m2l_mfma(
    const uint32_t *__restrict__ cls, ClassArgs ca, const uint32_t *__restrict__ occ, int level,
    const T2t<T> *__restrict__ mps, T2t<T> *__restrict__ mpr,
    const wbfmm_m2l_plan_t *__restrict__ P, const T *__restrict__ A,
    const T *__restrict__ A2, const T *__restrict__ CS,
    const T2t<T> *__restrict__ zero, int nsplit, T2t<T> *__restrict__ part, unsigned ntot) {
  using SH = M2LShape<N>;
  constexpr int NB = SH::NB, NU = SH::NU, S = SH::S;
  __shared__ T R[(M2LDouble<T, N>::value ? 2 : 1) * (S + 1) * 16];
  __shared__ int4 tmap1[4 * NB], tmap2[4 * NB], tmap3[4 * NB];
  __shared__ int4 uinf[NU];
  __shared__ T2t<T> cs[2][N + 1];
  __shared__ unsigned vm[8][7];
  __shared__ unsigned anym[7];

  // wv through readfirstlane: a scalar value, so the branch into the
  // per-wavefront code below is a scalar branch
  const int tid = threadIdx.x, lane = tid & 63, col = lane & 15, grp = lane >> 4;
  const int wv = __builtin_amdgcn_readfirstlane(tid >> 6);
  for (int i = tid; i < 4 * NB; i += 64 * W) {
    const int b = i >> 2, g = i & 3;
    int t1[4], t2[4], t3[4];
    for (int q = 0; q < 4; q++) {
      const int row = 16 * b + MF<T>::drow(g, q);
      t1[q] = P->map1[row] >> 6;
      t2[q] = P->map2[row] >> 6;
      t3[q] = P->map3[row] >> 6;
    }
    tmap1[i] = make_int4(t1[0], t1[1], t1[2], t1[3]);
    tmap2[i] = make_int4(t2[0], t2[1], t2[2], t2[3]);
    tmap3[i] = make_int4(t3[0], t3[1], t3[2], t3[3]);
  }
  // unit u = (n, m): index of x_n^0, m, the slots of P_m and Q_m
  for (int u = tid; u < NU; u += 64 * W) {
    int n = 0;
    while ((n + 1) * (n + 2) / 2 <= u) n++;
    const int m = u - n * (n + 1) / 2;
    uinf[u] = make_int4(n * (n + 1), m, P->s1[u] & 0xffff, (P->s1[u] >> 16) & 0xffff);
  }
  for (int i = tid; i < (M2LDouble<T, N>::value ? 2 : 1) * (S + 1) * 16; i += 64 * W)
    R[i] = T(0);
  for (int i = tid; i < 8 * 7; i += 64 * W) vm[i / 7][i % 7] = 0u;

  const int split = (int)(blockIdx.x % (unsigned)nsplit);
  const unsigned nch = ca.chunk_prefix[8], bid = blockIdx.x / (unsigned)nsplit, q8 = nch / 8;
  const unsigned chunk = bid < 8 * q8 ? (bid & 7) * q8 + (bid >> 3) : bid;
  int c = 0;
  while (c < 7 && chunk >= ca.chunk_prefix[c + 1]) c++;
  const int tg = col >> 1;
  const unsigned k8 = (chunk - ca.chunk_prefix[c]) * 8 + tg;
  const bool tvalid = k8 < ca.count[c];
  const uint32_t tb = tvalid ? cls[ca.start[c] + k8] : 0u;
  const int ti = compact3(tb), tj = compact3(tb >> 1), tk = compact3(tb >> 2);
  const int px = c & 1, py = (c >> 1) & 1, pz = (c >> 2) & 1;
  const int nbox = 1 << level;
  __syncthreads();

  // occupied list-4 sources per target (the real lane of each target)
  if ((col & 1) == 0) {
    for (int k = grp + 4 * wv; k < 216; k += 4 * W) {
      const int dx = k % 6 - 2 - px, dy = (k / 6) % 6 - 2 - py, dz = k / 36 - 2 - pz;
      if (dx >= -1 && dx <= 1 && dy >= -1 && dy <= 1 && dz >= -1 && dz <= 1) continue;
      const int si = ti + dx, sj = tj + dy, sk = tk + dz;
      if (!tvalid || si < 0 || si >= nbox || sj < 0 || sj >= nbox || sk < 0 || sk >= nbox)
        continue;
      const uint64_t sb = spread3(si) | (spread3(sj) << 1) | (spread3(sk) << 2);
      if ((occ[sb >> 5] >> (sb & 31)) & 1u) atomicOr(&vm[tg][k >> 5], 1u << (k & 31));
    }
  }
  __syncthreads();
  if (tid < 7) {
    unsigned a = 0;
    for (int j = 0; j < 8; j++) a |= vm[j][tid];
    anym[tid] = a;
  }
  __syncthreads();

#define HEL_LOOP(WI)                                                                          \
  hel_m2l_loop<T, N, W, WI>(R, tmap1, tmap2, tmap3, uinf, cs, vm, anym, lane, col, grp, tvalid,  \
                         tb, ti, tj, tk, px, py, pz, c, k8, split, nsplit, ca, P, A, A2, CS, \
                         zero, mps, mpr, part, ntot)
  static_assert(W == 4 || W == 8, "four or eight wavefronts per block");
  if constexpr (W == 4) {
    switch (wv) {
      case 0: HEL_LOOP(0); break;
      case 1: HEL_LOOP(1); break;
      case 2: HEL_LOOP(2); break;
      default: HEL_LOOP(3); break;
    }
  } else {
    switch (wv) {
      case 0: HEL_LOOP(0); break;
      case 1: HEL_LOOP(1); break;
      case 2: HEL_LOOP(2); break;
      case 3: HEL_LOOP(3); break;
      case 4: HEL_LOOP(4); break;
      case 5: HEL_LOOP(5); break;
      case 6: HEL_LOOP(6); break;
      default: HEL_LOOP(7); break;
    }
  }
#undef HEL_LOOP
}

// ---------------------------------------------------------------------
// Sparse levels with the offsets packed by class (FP32 levels).  A block
// takes 16 targets of one Morton class.  For each offset class (offsets
// sharing dz and dx^2 + dy^2, hence all three stage tables) it lists the
// (target, member) pairs with an occupied source in target-major order and
// deals them round-robin over T = max(ceil(n / 8), most pairs of one
// target) tiles of 8 pairs: position p goes to tile p mod T, column pair
// p div T (columns 2q, 2q + 1 are the pair's real and imaginary lanes, as
// in the chunked kernel).  A target's pairs are consecutive positions, at
// most T of them, so a tile holds each target at most once.  The chunked
// kernel runs the union of its 8 targets' offsets; on the 1e6-point sphere
// half of that work has no source.  The local expansions stay in
// registers on owner lanes: lane L = 64 w + lane holds L^{+nu} (sign 0) or
// L^{-nu} (sign 1) of target L & 15, sign (L >> 4) & 1, for the units
// (L >> 5) + 2W i; after each tile's third stage it reads its target's
// column pair, found from (r, q, cnt) = (s mod T, s div T, count), s the
// target's first position.  Each target sums its translations in tile
// order.
// ---------------------------------------------------------------------
constexpr int kHelPackT = 16;

template <typename T, int N>
struct HelPackedFits {
  static constexpr bool value = sizeof(T) == 4 && N >= 8 && N <= 27;
};

template <typename T, int N, int W, int WI>
__device__ __forceinline__ void hel_packed_loop(
    T *R, const int4 *tmap1, const int4 *tmap2, const int4 *tmap3, const int4 *uinf,
    T2t<T> (*cs)[8][N + 1], int ncls, const uint16_t *gbase, const uint16_t *gcount,
    const uint8_t *gT, const int2 *gtab, const uint8_t (*gmem)[8], const uint8_t *gnm,
    const uint8_t *pairs, const uint16_t (*zinfo)[kHelPackT], const int4 *ktab,
    const int4 *tco, const uint32_t *tbox, int lane, int col, int grp,
    const T *__restrict__ A, const T *__restrict__ A2, const T *__restrict__ CS,
    const T2t<T> *__restrict__ zero, const T2t<T> *__restrict__ mps,
    T2t<T> *__restrict__ mpr) {
  using SH = M2LShape<N>;
  constexpr int C = SH::C, NU = SH::NU, KT = SH::KT;
  constexpr int NAW = (NU + 4 * W - 1) / (4 * W);  // units per lane (deposit)
  constexpr WaveSeq WS = make_seq(SH::BN, W, WI);
  constexpr int NBW = WS.nb;
  constexpr bool DB = M2LDouble<T, N>::value;
  constexpr int G = 2 * W;                       // unit groups of the owners
  constexpr int NAQ = (NU + G - 1) / G;          // owner units per lane
  T *R1 = R + (SH::S + 1) * 16;  // the second buffer (DB only)
  const int ubase = grp + 4 * WI;
  const int LL = 64 * WI + lane, ot = LL & 15, osg = (LL >> 4) & 1, og = LL >> 5;

  auto next_tile = [&](int g, int t, int &gn, int &tn) {
    if (g >= 0 && g < ncls && t + 1 < gT[g]) {
      gn = g;
      tn = t + 1;
      return;
    }
    for (g++; g < ncls; g++)
      if (gcount[g]) break;
    gn = g;
    tn = 0;
  };
  // this lane's column pair of tile (g, t): its source expansion and
  // member index (zeros and member 0 for an empty pair); with fill, the
  // class's phases into cs[p] (every wavefront writes the same values)
  auto source_of = [&](int g, int t, int p, bool fill, int &icol) -> const T2t<T> * {
    const int nm = gnm[g];
    if (fill)
      for (int e = lane; e < nm * (N + 1); e += 64) {
        const int i = e / (N + 1), m = e - i * (N + 1);
        const T *tt = CS + ((size_t)ktab[gmem[g][i]].w * (N + 1) + m) * 2;
        cs[p][i][m] = mk2<T>(tt[0], tt[1]);
      }
    const int Tt = gT[g], pos = t + (col >> 1) * Tt;
    icol = 0;
    if (pos >= gcount[g]) return zero;
    const int pr = pairs[gbase[g] + pos];
    icol = pr >> 4;
    const int4 kt = ktab[gmem[g][icol]];
    const uint32_t sb = morton_add(tbox[pr & 15], (uint32_t)kt.x, (uint32_t)kt.y, (uint32_t)kt.z);
    return mps + (uint64_t)sb * C;
  };
  T pre_re[NAW], pre_im[NAW];
  // x_{+m} on the real lane, x_{-m} on the imaginary one
  auto load_sources = [&](const T2t<T> *src, const int4 *tu, int ub, int sp) {
#pragma unroll
    for (int i = 0; i < NAW; i++) {
      const int u = min(ub + 4 * W * i, NU - 1);
      const int4 inf = tu[u];
      const T2t<T> q = src[inf.x + (sp ? -inf.y : inf.y)];
      pre_re[i] = q.x;
      pre_im[i] = q.y;
    }
  };
  // x'_{+-m} = e^{+-i m chi} x_{+-m}; the partner lane holds the other sign
  auto deposit = [&](const int4 *tu, int ub, int cl, int p, int ic) {
    const int sp = cl & 1;
#pragma unroll
    for (int i = 0; i < NAW; i++) {
      const int u = min(ub + 4 * W * i, NU - 1);
      const int4 inf = tu[u];
      const T2t<T> e = cs[p][ic][inf.y];
      const T es = sp ? -e.y : e.y;
      const T xr = e.x * pre_re[i] - es * pre_im[i];
      const T xi = es * pre_re[i] + e.x * pre_im[i];
      const T yr = __shfl_xor(xr, 1), yi = __shfl_xor(xi, 1);
      if (ub + 4 * W * i >= NU) continue;
      R[inf.z * 16 + cl] = sp ? yi + xi : xr + yr;
      R[inf.w * 16 + cl] = sp ? yi - xi : xr - yr;
    }
  };
  auto write_out = [&](const int4 *tm, const V4t<T> *y, int gp, int cl) {
#pragma unroll
    for (int lb = 0; lb < NBW; lb++) {
      const int4 d = tm[4 * (WI + W * lb) + gp];
      R[d.x * 16 + cl] = y[lb][0];
      R[d.y * 16 + cl] = y[lb][1];
      R[d.z * 16 + cl] = y[lb][2];
      R[d.w * 16 + cl] = y[lb][3];
    }
  };

  T acc_re[NAQ], acc_im[NAQ];
#pragma unroll
  for (int i = 0; i < NAQ; i++) acc_re[i] = acc_im[i] = T(0);

  int g, t, p = 0, icol = 0, icol_n = 0;
  next_tile(-1, 0, g, t);
  g = __builtin_amdgcn_readfirstlane(g);
  int th = 0, ri = 0;
  if (g < ncls) {
    const int2 tr = gtab[g];
    th = __builtin_amdgcn_readfirstlane(tr.x);
    ri = __builtin_amdgcn_readfirstlane(tr.y);
    load_sources(source_of(g, t, 0, true, icol), uinf, ubase, col & 1);
    deposit(uinf, ubase, col, 0, icol);
  }
  __syncthreads();
  while (g < ncls) {
    int z0;
    asm volatile("s_mov_b32 %0, 0" : "=s"(z0));
    const int4 *t1 = tmap1 + z0, *t2 = tmap2 + z0, *t3 = tmap3 + z0, *tu = uinf + z0;
    const int ln = lane + z0, cl = ln & 15, gp = ln >> 4, ub = gp + 4 * WI, sp = cl & 1;
    int gn, tn;
    next_tile(g, t, gn, tn);
    gn = __builtin_amdgcn_readfirstlane(gn);
    tn = __builtin_amdgcn_readfirstlane(tn);
    // the phase buffer changes with the class: a new class's phases go
    // into cs[p ^ 1], last read (by the previous class) before the barrier
    // that ended the previous tile; the tiles of one class share cs[p]
    const bool newc = gn != g;
    const int pn = newc ? p ^ 1 : p;
    int th_n = 0, ri_n = 0;
    if (gn < ncls) {
      const int2 tr = gtab[gn];
      th_n = __builtin_amdgcn_readfirstlane(tr.x);
      ri_n = __builtin_amdgcn_readfirstlane(tr.y);
    }
    // the next tile's sources are requested after the first stage, as in
    // the chunked kernel (their registers are not held across all three)
    T *RZ = R;
    if constexpr (DB) {
      stage_wo<T, N, W, WI>(A + (size_t)th * KT * 64, R, R1, t1, ln, gp, cl);
      __syncthreads();
      if (gn < ncls) load_sources(source_of(gn, tn, pn, newc, icol_n), tu, ub, sp);
      stage_cx_wo<T, N, W, WI>(A2 + (size_t)(2 * ri) * KT * 64, A2 + (size_t)(2 * ri + 1) * KT * 64,
                               R1, R, t2, ln, gp, cl);
      __syncthreads();
      stage_wo<T, N, W, WI>(A + (size_t)th * KT * 64, R, R1, t3, ln, gp, cl);
      __syncthreads();
      RZ = R1;
    } else {
      V4t<T> y[NBW > 0 ? NBW : 1];
      stage_w<T, N, W, WI>(A + (size_t)th * KT * 64, R, ln, gp, cl, y);
      __syncthreads();
      write_out(t1, y, gp, cl);
      __syncthreads();
      if (gn < ncls) load_sources(source_of(gn, tn, pn, newc, icol_n), tu, ub, sp);
      stage_cx<T, N, W, WI>(A2 + (size_t)(2 * ri) * KT * 64, A2 + (size_t)(2 * ri + 1) * KT * 64,
                            R, ln, gp, cl, y);
      __syncthreads();
      write_out(t2, y, gp, cl);
      __syncthreads();
      stage_w<T, N, W, WI>(A + (size_t)th * KT * 64, R, ln, gp, cl, y);
      __syncthreads();
      write_out(t3, y, gp, cl);
      __syncthreads();
    }
    // owners: y_{+-nu} = (P +- Q)/2 of the target's column pair, then
    // L^{+-nu} += e^{-+i nu chi} y
    {
      const int Tt = gT[g], z = zinfo[g][ot];
      const int r = z & 15, q = (z >> 4) & 15, cnt = z >> 8;
      int d = t - r;
      if (d < 0) d += Tt;
      if (d < cnt) {
        const int cc = q + (r + d >= Tt ? 1 : 0);
        const int ic = pairs[gbase[g] + t + cc * Tt] >> 4;
        const int c0 = 2 * cc;
#pragma unroll
        for (int i = 0; i < NAQ; i++) {
          const int u = og + G * i;
          if (u >= NU) continue;
          const int nu = tu[u].y;
          const T2t<T> e = cs[p][ic][nu];
          const T es = osg ? -e.y : e.y;
          const T pr = RZ[u * 16 + c0], pi = RZ[u * 16 + c0 + 1];
          const T qr = nu ? RZ[(NU + u) * 16 + c0] : T(0);
          const T qi = nu ? RZ[(NU + u) * 16 + c0 + 1] : T(0);
          const T yr = T(0.5) * (osg ? pr - qr : pr + qr);
          const T yi = T(0.5) * (osg ? pi - qi : pi + qi);
          acc_re[i] += e.x * yr + es * yi;
          acc_im[i] += e.x * yi - es * yr;
        }
      }
    }
    if constexpr (!DB) __syncthreads();  // every read of R done before the next deposit
    if (gn < ncls) deposit(tu, ub, cl, pn, icol_n);
    __syncthreads();
    p = pn;
    g = gn;
    t = tn;
    th = th_n;
    ri = ri_n;
    icol = icol_n;
  }
  (void)icol;
  if (!tco[ot].w) return;
  T2t<T> *dst = mpr + (size_t)tbox[ot] * C;
#pragma unroll
  for (int i = 0; i < NAQ; i++) {
    const int u = og + G * i;
    if (u >= NU) continue;
    const int4 inf = uinf[u];
    if (osg && inf.y == 0) continue;  // L^0 is the sign-0 owner's
    const int r = inf.x + (osg ? -inf.y : inf.y);
    const T2t<T> o = dst[r];
    dst[r] = mk2<T>(o.x + acc_re[i], o.y + acc_im[i]);
  }
}

// sixteen wavefronts from order 16 (one block per CU, four wavefronts per
// SIMD at 128 registers); eight below, two blocks per CU
template <typename T, int N, int W>
__global__ __launch_bounds__(64 * W) __attribute__((amdgpu_waves_per_eu(
    W == 16 ? 4 : hel_m2l_waves_per_eu<T, N, W>()))) void k_hel_m2l_packed(
    const uint32_t *__restrict__ cls, ClassArgs ca, const uint32_t *__restrict__ occ, int level,
    const T2t<T> *__restrict__ mps, T2t<T> *__restrict__ mpr,
    const wbfmm_m2l_plan_t *__restrict__ P, const T *__restrict__ A,
    const T *__restrict__ A2, const T *__restrict__ CS, const T2t<T> *__restrict__ zero) {
  using SH = M2LShape<N>;
  constexpr int NB = SH::NB, NU = SH::NU, S = SH::S;
  static_assert(W == 4 || W == 8 || W == 16, "four, eight or sixteen wavefronts per block");
  __shared__ T R[(M2LDouble<T, N>::value ? 2 : 1) * (S + 1) * 16];
  __shared__ int4 tmap1[4 * NB], tmap2[4 * NB], tmap3[4 * NB];
  __shared__ int4 uinf[NU];
  __shared__ T2t<T> cs[2][8][N + 1];
  __shared__ unsigned vm[kHelPackT][7];
  __shared__ uint8_t pairs[kHelPackT * 189];
  __shared__ uint16_t zinfo[kMaxOffCls][kHelPackT];  // r | q << 4 | cnt << 8
  __shared__ uint8_t ccnt[kMaxOffCls][kHelPackT];
  __shared__ uint16_t gbase[kMaxOffCls], gcount[kMaxOffCls];
  __shared__ uint8_t gT[kMaxOffCls], gnm[kMaxOffCls], gmem[kMaxOffCls][8];
  __shared__ int2 gtab[kMaxOffCls];
  __shared__ int4 ktab[216];  // dilated offsets (x, y, z) and offset index
  __shared__ int4 tco[kHelPackT];
  __shared__ uint32_t tbox[kHelPackT];

  const int tid = threadIdx.x, lane = tid & 63, col = lane & 15, grp = lane >> 4;
  const int wv = __builtin_amdgcn_readfirstlane(tid >> 6);
  for (int i = tid; i < 4 * NB; i += 64 * W) {
    const int b = i >> 2, g = i & 3;
    int t1[4], t2[4], t3[4];
    for (int q = 0; q < 4; q++) {
      const int row = 16 * b + MF<T>::drow(g, q);
      t1[q] = P->map1[row] >> 6;
      t2[q] = P->map2[row] >> 6;
      t3[q] = P->map3[row] >> 6;
    }
    tmap1[i] = make_int4(t1[0], t1[1], t1[2], t1[3]);
    tmap2[i] = make_int4(t2[0], t2[1], t2[2], t2[3]);
    tmap3[i] = make_int4(t3[0], t3[1], t3[2], t3[3]);
  }
  // unit u = (n, m): index of x_n^0, m, the slots of P_m and Q_m
  for (int u = tid; u < NU; u += 64 * W) {
    int n = 0;
    while ((n + 1) * (n + 2) / 2 <= u) n++;
    const int m = u - n * (n + 1) / 2;
    uinf[u] = make_int4(n * (n + 1), m, P->s1[u] & 0xffff, (P->s1[u] >> 16) & 0xffff);
  }
  for (int i = tid; i < (M2LDouble<T, N>::value ? 2 : 1) * (S + 1) * 16; i += 64 * W)
    R[i] = T(0);
  for (int i = tid; i < kHelPackT * 7; i += 64 * W) vm[i / 7][i % 7] = 0u;

  int c;
  unsigned cj;
  // a block with no chunk leaves before any barrier (uniform over the block)
  if (!block_chunk(blockIdx.x, ca, c, cj)) return;
  const int px = c & 1, py = (c >> 1) & 1, pz = (c >> 2) & 1;
  const int nbox = 1 << level;
  const int ncls = kOffCls[c].n;
  for (int k = tid; k < 216; k += 64 * W) {
    const int dx = k % 6 - 2 - px, dy = (k / 6) % 6 - 2 - py, dz = k / 36 - 2 - pz;
    const int off = (dx + 3) * 49 + (dy + 3) * 7 + dz + 3;
    ktab[k] = make_int4((int)dilate_offset(dx, 0, kMortonX), (int)dilate_offset(dy, 1, kMortonY),
                        (int)dilate_offset(dz, 2, kMortonZ), off);
  }
  for (int g = tid; g < ncls; g += 64 * W) {
    gnm[g] = kOffCls[c].nm[g];
    for (int i = 0; i < 8; i++) gmem[g][i] = kOffCls[c].k[g][i];
    // the members share the polar angle and distance: the first one's
    const int k0 = kOffCls[c].k[g][0];
    const int dx = k0 % 6 - 2 - px, dy = (k0 / 6) % 6 - 2 - py, dz = k0 / 36 - 2 - pz;
    const int off = (dx + 3) * 49 + (dy + 3) * 7 + dz + 3;
    gtab[g] = make_int2(P->th[off], P->rix[off]);
  }
  if (tid < kHelPackT) {
    const unsigned kk = cj * kHelPackT + tid;
    const bool tv = kk < ca.count[c];
    const uint32_t tb = tv ? cls[ca.start[c] + kk] : 0u;
    tbox[tid] = tb;
    tco[tid] = make_int4(compact3(tb), compact3(tb >> 1), compact3(tb >> 2), tv ? 1 : 0);
  }
  __syncthreads();
  for (int idx = tid; idx < kHelPackT * 216; idx += 64 * W) {
    const int j = idx % kHelPackT, k = idx / kHelPackT;
    const int dx = k % 6 - 2 - px, dy = (k / 6) % 6 - 2 - py, dz = k / 36 - 2 - pz;
    if (dx >= -1 && dx <= 1 && dy >= -1 && dy <= 1 && dz >= -1 && dz <= 1) continue;
    const int4 tc = tco[j];
    const int si = tc.x + dx, sj = tc.y + dy, sk = tc.z + dz;
    if (!tc.w || si < 0 || si >= nbox || sj < 0 || sj >= nbox || sk < 0 || sk >= nbox) continue;
    const uint64_t sb = spread3(si) | (spread3(sj) << 1) | (spread3(sk) << 2);
    if ((occ[sb >> 5] >> (sb & 31)) & 1u) atomicOr(&vm[j][k >> 5], 1u << (k & 31));
  }
  __syncthreads();
  for (int idx = tid; idx < ncls * kHelPackT; idx += 64 * W) {
    const int g = idx / kHelPackT, j = idx % kHelPackT;
    int n = 0;
    for (int i = 0; i < gnm[g]; i++) {
      const int k = gmem[g][i];
      n += (vm[j][k >> 5] >> (k & 31)) & 1u;
    }
    ccnt[g][j] = (uint8_t)n;
  }
  __syncthreads();
  // per class: pair count, tiles of 8 pairs, each target's (r, q, cnt)
  for (int g = tid; g < ncls; g += 64 * W) {
    int n = 0, mx = 0;
    for (int j = 0; j < kHelPackT; j++) {
      n += ccnt[g][j];
      mx = max(mx, (int)ccnt[g][j]);
    }
    const int Tt = n ? max((n + 7) / 8, mx) : 0;
    gcount[g] = (uint16_t)n;
    gT[g] = (uint8_t)Tt;
    int s = 0;
    for (int j = 0; j < kHelPackT; j++) {
      const int cn = ccnt[g][j];
      // (a target without pairs may start at s = n = 8 T: stored as 0)
      zinfo[g][j] = (uint16_t)(cn ? ((s % Tt) | ((s / Tt) << 4) | (cn << 8)) : 0);
      s += cn;
    }
  }
  __syncthreads();
  if (tid == 0) {
    int b = 0;
    for (int g = 0; g < ncls; g++) {
      gbase[g] = (uint16_t)b;
      b += gcount[g];
    }
  }
  __syncthreads();
  // pairs (target | member << 4), target-major within each class
  for (int idx = tid; idx < ncls * kHelPackT; idx += 64 * W) {
    const int g = idx / kHelPackT, j = idx % kHelPackT;
    if (!ccnt[g][j]) continue;
    int s = gbase[g];
    for (int jj = 0; jj < j; jj++) s += ccnt[g][jj];
    for (int i = 0; i < gnm[g]; i++) {
      const int k = gmem[g][i];
      if ((vm[j][k >> 5] >> (k & 31)) & 1u) pairs[s++] = (uint8_t)(j | (i << 4));
    }
  }
  __syncthreads();

#define HEL_PACKED(WI)                                                                       \
  hel_packed_loop<T, N, W, WI>(R, tmap1, tmap2, tmap3, uinf, cs, ncls, gbase, gcount, gT,    \
                               gtab, gmem, gnm, pairs, zinfo, ktab, tco, tbox, lane, col, grp, \
                               A, A2, CS, zero, mps, mpr)
  if constexpr (W == 4) {
    switch (wv) {
      case 0: HEL_PACKED(0); break;
      case 1: HEL_PACKED(1); break;
      case 2: HEL_PACKED(2); break;
      default: HEL_PACKED(3); break;
    }
  } else if constexpr (W == 8) {
    switch (wv) {
      case 0: HEL_PACKED(0); break;
      case 1: HEL_PACKED(1); break;
      case 2: HEL_PACKED(2); break;
      case 3: HEL_PACKED(3); break;
      case 4: HEL_PACKED(4); break;
      case 5: HEL_PACKED(5); break;
      case 6: HEL_PACKED(6); break;
      default: HEL_PACKED(7); break;
    }
  } else {
    switch (wv) {
      case 0: HEL_PACKED(0); break;
      case 1: HEL_PACKED(1); break;
      case 2: HEL_PACKED(2); break;
      case 3: HEL_PACKED(3); break;
      case 4: HEL_PACKED(4); break;
      case 5: HEL_PACKED(5); break;
      case 6: HEL_PACKED(6); break;
      case 7: HEL_PACKED(7); break;
      case 8: HEL_PACKED(8); break;
      case 9: HEL_PACKED(9); break;
      case 10: HEL_PACKED(10); break;
      case 11: HEL_PACKED(11); break;
      case 12: HEL_PACKED(12); break;
      case 13: HEL_PACKED(13); break;
      case 14: HEL_PACKED(14); break;
      default: HEL_PACKED(15); break;
    }
  }
#undef HEL_PACKED
}

// the packed kernel (FP32 levels) where fewer than 60 % of the level's
// boxes hold targets and there are at least 512 chunks of 16 targets;
// WBFMM_HEL_PACKED=0 or 1 forces the choice at every level and size
static int hel_packed_choice() {
  const char *e = getenv("WBFMM_HEL_PACKED");
  return e ? atoi(e) : -1;
}

template <typename T, int N>
static int launch_hel_m2l(const uint32_t *cls, const ClassArgs &ca, unsigned nchunks,
                          const uint32_t *occ, int level, const T *mps, T *mpr,
                          const void *plan, const T *A, const T *A2,
                          const T *CS, const T *zero,
                          const wbfmm_m2l_plan_t *hplan) {
  if (!plan_matches<N>(hplan)) {
    fprintf(stderr, "wbfmm (HIP): M2L plan for order %d does not match the kernel's bins\n", N);
    return WBFMM_HIP_ERR_ARG;
  }
  // the offsets of a chunk split over up to 8 blocks when there are few
  // chunks (the split depends on the chunk count only: fixed summation
  // order for a given problem)
  constexpr int C = (N + 1) * (N + 1);
  unsigned ntot = 0;
  for (int c = 0; c < 8; c++) ntot += ca.count[c];
  if constexpr (HelPackedFits<T, N>::value) {
    ClassArgs cp = ca;
    unsigned nch = 0;
    for (int c = 0; c < 8; c++) {
      cp.chunk_prefix[c] = nch;
      nch += (ca.count[c] + kHelPackT - 1) / kHelPackT;
    }
    cp.chunk_prefix[8] = nch;
    cp.qc = 0;
    // the calling thread's variant (wbfmm_hip_set_variant): 0 chunks at
    // the sparse levels, 1 or 2 the packed kernel there
    const int var = wbfmm_hip_get_variant(WBFMM_VARIANT_M2L);
    const int pc = var == 0 ? 0 : hel_packed_choice();
    const double fill = (double)ntot / (double)((uint64_t)1 << (3 * level));
    if (nch > 0 && (pc == 1 || (pc == -1 && nch >= 512 && fill < 0.6))) {
      constexpr int W = N >= 16 ? 16 : 4;
      WBFMM_LAUNCH("hel_m2l_f32", (k_hel_m2l_packed<T, N, W>), dim3(nch), dim3(64 * W), 0, cls,
                   cp, occ, level, (const T2t<T> *)mps, (T2t<T> *)mpr,
                   (const wbfmm_m2l_plan_t *)plan, A, A2, CS, (const T2t<T> *)zero);
      return 0;
    }
  }
  const int nsplit = nchunks >= 1024 ? 1 : (int)(2048 / nchunks < 8 ? 2048 / nchunks : 8);
  T2t<T> *part = nullptr;
  if (nsplit > 1 && dmalloc((void **)&part, sizeof(T2t<T>) * (size_t)nsplit * ntot * C))
    return WBFMM_HIP_ERR_NOMEM;
  // eight wavefronts per block from order 16, where four would need more
  // than 256 registers each (the LDS buffer allows one block per CU there)
  constexpr int W = N >= 16 ? 8 : 4;
  WBFMM_LAUNCH(sizeof(T) == 4 ? "hel_m2l_f32" : "hel_m2l", (k_hel_m2l_mfma<T, N, W>),
               dim3(nchunks * nsplit), dim3(64 * W), 0, cls, ca,
               occ, level, (const T2t<T> *)mps, (T2t<T> *)mpr,
               (const wbfmm_m2l_plan_t *)plan, A, A2, CS, (const T2t<T> *)zero, nsplit, part,
               ntot);
  if (nsplit > 1) {
    const size_t ne = (size_t)ntot * 2 * C;
    WBFMM_LAUNCH("m2l_reduce", k_m2l_reduce<T>, dim3((unsigned)((ne + 255) / 256)),
                 dim3(256), 0, cls, ntot, nsplit, 2 * C, (const T *)part, mpr);
    dfree(part);
  }
  return 0;
}


// ---------------------------------------------------------------------
// M2M (MODE 0, src/pass.c:581-632, src/shift.c:244-401) and L2L (MODE 1,
// src/pass.c:550-576, src/shift.c:403-585) in the same form: the rotate /
// coaxial R|R / rotate-back of each child direction c, whose polar angle
// is acos(1/sqrt 3) (tile 0) or acos(-1/sqrt 3) (tile 1) and whose phases
// are CS[c] (cos, sin of m chi_c).  The R|R table is real: stage 2 takes
// one product per tile step.  A block of W wavefronts takes 8 parents
// (columns 2t + s as in the M2L).  M2M adds the 8 children of each parent
// in registers, in child order, and writes the parent; L2L translates each
// parent's local expansion to each flagged child and adds it there.
// ---------------------------------------------------------------------
template <typename T, int N, int W, int WI, int MODE>
__device__ __forceinline__ void hel_shift_loop(
    T *R, const int4 *tmap1, const int4 *tmap2, const int4 *tmap3, const int4 *uinf,
    const T2t<T> (*cs)[N + 1], const unsigned *cmask, unsigned anyc, int lane, int col, int grp,
    bool pvalid, uint64_t p, const T2t<T> *__restrict__ in, T2t<T> *__restrict__ out,
    const T *__restrict__ A, const T *__restrict__ A2, const T *__restrict__ CS,
    const T2t<T> *__restrict__ zero) {
  using SH = M2LShape<N>;
  constexpr int C = SH::C, NU = SH::NU, KT = SH::KT;
  constexpr int NAW = (NU + 4 * W - 1) / (4 * W);
  constexpr WaveSeq WS = make_seq(SH::BN, W, WI);
  constexpr int NBW = WS.nb;
  const int ubase = grp + 4 * WI, tg = col >> 1;

  T pre_re[NAW], pre_im[NAW];
  auto load_sources = [&](const T2t<T> *src, const int4 *tu, int ub, int sp) {
#pragma unroll
    for (int i = 0; i < NAW; i++) {
      const int u = min(ub + 4 * W * i, NU - 1);
      const int4 inf = tu[u];
      const T2t<T> q = src[inf.x + (sp ? -inf.y : inf.y)];
      pre_re[i] = q.x;
      pre_im[i] = q.y;
    }
  };
  auto deposit = [&](const int4 *tu, int ub, int cl, int c) {
    const int sp = cl & 1;
#pragma unroll
    for (int i = 0; i < NAW; i++) {
      const int u = min(ub + 4 * W * i, NU - 1);
      const int4 inf = tu[u];
      const T2t<T> e = cs[c][inf.y];
      const T es = sp ? -e.y : e.y;
      const T xr = e.x * pre_re[i] - es * pre_im[i];
      const T xi = es * pre_re[i] + e.x * pre_im[i];
      const T yr = __shfl_xor(xr, 1), yi = __shfl_xor(xi, 1);
      if (ub + 4 * W * i >= NU) continue;
      R[inf.z * 16 + cl] = sp ? yi + xi : xr + yr;
      R[inf.w * 16 + cl] = sp ? yi - xi : xr - yr;
    }
  };
  auto write_out = [&](const int4 *tm, const V4t<T> *y, int gp, int cl) {
#pragma unroll
    for (int lb = 0; lb < NBW; lb++) {
      const int4 d = tm[4 * (WI + W * lb) + gp];
      R[d.x * 16 + cl] = y[lb][0];
      R[d.y * 16 + cl] = y[lb][1];
      R[d.z * 16 + cl] = y[lb][2];
      R[d.w * 16 + cl] = y[lb][3];
    }
  };

  T acc_re[NAW], acc_im[NAW];
#pragma unroll
  for (int i = 0; i < NAW; i++) acc_re[i] = acc_im[i] = T(0);
  if (MODE == 1) load_sources(pvalid ? in + p * (uint64_t)C : zero, uinf, ubase, col & 1);

  for (int c = 0; c < 8; c++) {
    if (!((anyc >> c) & 1u)) continue;  // the same in every lane of the block
    int z0;
    asm volatile("s_mov_b32 %0, 0" : "=s"(z0));
    const int4 *t1 = tmap1 + z0, *t2 = tmap2 + z0, *t3 = tmap3 + z0, *tu = uinf + z0;
    const int ln = lane + z0, cl = ln & 15, gp = ln >> 4, ub = gp + 4 * WI, sp = cl & 1;
    const bool v = (cmask[tg] >> c) & 1u;
    const int th = MODE == 0 ? (c >> 2) & 1 : 1 - ((c >> 2) & 1);
    if (MODE == 0) load_sources(v ? in + (8 * p + c) * (uint64_t)C : zero, tu, ub, sp);
    __syncthreads();  // every read of R by the last direction done
    deposit(tu, ub, cl, c);
    __syncthreads();

    V4t<T> y[NBW > 0 ? NBW : 1];
    stage_w<T, N, W, WI>(A + (size_t)th * KT * 64, R, ln, gp, cl, y);
    __syncthreads();
    write_out(t1, y, gp, cl);
    __syncthreads();
    stage_w<T, N, W, WI>(A2, R, ln, gp, cl, y);
    __syncthreads();
    write_out(t2, y, gp, cl);
    __syncthreads();
    stage_w<T, N, W, WI>(A + (size_t)th * KT * 64, R, ln, gp, cl, y);
    __syncthreads();
    write_out(t3, y, gp, cl);
    __syncthreads();

    const int c0 = cl & ~1;
    T2t<T> *dst = out + (8 * p + c) * (uint64_t)C;
#pragma unroll
    for (int i = 0; i < NAW; i++) {
      const int u = min(ub + 4 * W * i, NU - 1);
      const int4 inf = tu[u];
      const int nu = inf.y;
      const T2t<T> e = cs[c][nu];
      const T es = sp ? -e.y : e.y;
      const T pr = R[u * 16 + c0], pi = R[u * 16 + c0 + 1];
      const T qr = nu ? R[(NU + u) * 16 + c0] : T(0);
      const T qi = nu ? R[(NU + u) * 16 + c0 + 1] : T(0);
      const T yr = T(0.5) * (sp ? pr - qr : pr + qr);
      const T yi = T(0.5) * (sp ? pi - qi : pi + qi);
      const T orr = e.x * yr + es * yi, oi = e.x * yi - es * yr;
      if (MODE == 0) {
        acc_re[i] += orr;
        acc_im[i] += oi;
      } else if (v && ub + 4 * W * i < NU && !(sp && nu == 0)) {
        const int r = inf.x + (sp ? -nu : nu);
        const T2t<T> o = dst[r];
        dst[r] = mk2<T>(o.x + orr, o.y + oi);
      }
    }
  }
  if (MODE == 0 && pvalid) {
    T2t<T> *dst = out + p * (uint64_t)C;
    const int sp = col & 1;
#pragma unroll
    for (int i = 0; i < NAW; i++) {
      const int u = ubase + 4 * W * i;
      if (u >= NU) continue;
      const int4 inf = uinf[u];
      if (sp && inf.y == 0) continue;
      dst[inf.x + (sp ? -inf.y : inf.y)] = mk2<T>(acc_re[i], acc_im[i]);
    }
  }
}

template <typename T, int N, int W, int MODE>
__global__ __launch_bounds__(64 * W) void k_hel_shift_mfma(
    const uint32_t *__restrict__ parents, unsigned npar, const uint32_t *__restrict__ cstart,
    const uint32_t *__restrict__ cend, const uint32_t *__restrict__ cflag,
    const T2t<T> *__restrict__ in, T2t<T> *__restrict__ out,
    const wbfmm_m2l_plan_t *__restrict__ P, const T *__restrict__ A,
    const T *__restrict__ A2, const T *__restrict__ CS,
    const T2t<T> *__restrict__ zero) {
  using SH = M2LShape<N>;
  constexpr int NB = SH::NB, NU = SH::NU, S = SH::S;
  __shared__ T R[(S + 1) * 16];
  __shared__ int4 tmap1[4 * NB], tmap2[4 * NB], tmap3[4 * NB];
  __shared__ int4 uinf[NU];
  __shared__ T2t<T> cs[8][N + 1];
  __shared__ unsigned cmask[8];
  __shared__ unsigned anyc;

  const int tid = threadIdx.x, lane = tid & 63, col = lane & 15, grp = lane >> 4;
  const int wv = __builtin_amdgcn_readfirstlane(tid >> 6);
  for (int i = tid; i < 4 * NB; i += 64 * W) {
    const int b = i >> 2, g = i & 3;
    int t1[4], t2[4], t3[4];
    for (int q = 0; q < 4; q++) {
      const int row = 16 * b + MF<T>::drow(g, q);
      t1[q] = P->map1[row] >> 6;
      t2[q] = P->map2[row] >> 6;
      t3[q] = P->map3[row] >> 6;
    }
    tmap1[i] = make_int4(t1[0], t1[1], t1[2], t1[3]);
    tmap2[i] = make_int4(t2[0], t2[1], t2[2], t2[3]);
    tmap3[i] = make_int4(t3[0], t3[1], t3[2], t3[3]);
  }
  for (int u = tid; u < NU; u += 64 * W) {
    int n = 0;
    while ((n + 1) * (n + 2) / 2 <= u) n++;
    const int m = u - n * (n + 1) / 2;
    uinf[u] = make_int4(n * (n + 1), m, P->s1[u] & 0xffff, (P->s1[u] >> 16) & 0xffff);
  }
  for (int i = tid; i < (S + 1) * 16; i += 64 * W) R[i] = T(0);
  for (int i = tid; i < 8 * (N + 1); i += 64 * W)
    cs[i / (N + 1)][i % (N + 1)] = mk2<T>(CS[2 * i], CS[2 * i + 1]);
  // children to translate, per parent column: occupied children (M2M) or
  // flagged ones (L2L)
  const unsigned pi = blockIdx.x * 8 + (col >> 1);
  const bool pvalid = pi < npar;
  const uint64_t p = pvalid ? parents[pi] : 0;
  if (tid < 8) {
    const unsigned pj = blockIdx.x * 8 + tid;
    unsigned m = 0;
    if (pj < npar) {
      const uint64_t pp = parents[pj];
      for (int c = 0; c < 8; c++) {
        const uint64_t cb = 8 * pp + c;
        const bool v = MODE == 0 ? cend[cb] > cstart[cb] : (!cflag || cflag[cb] != 0);
        if (v) m |= 1u << c;
      }
    }
    cmask[tid] = m;
  }
  __syncthreads();
  if (tid == 0) {
    unsigned a = 0;
    for (int j = 0; j < 8; j++) a |= cmask[j];
    anyc = a;
  }
  __syncthreads();
  const unsigned any = __builtin_amdgcn_readfirstlane(anyc);

#define HEL_SHIFT(WI)                                                                        \
  hel_shift_loop<T, N, W, WI, MODE>(R, tmap1, tmap2, tmap3, uinf, cs, cmask, any, lane, col,   \
                                 grp, pvalid, p, in, out, A, A2, CS, zero)
  if constexpr (W == 4) {
    switch (wv) {
      case 0: HEL_SHIFT(0); break;
      case 1: HEL_SHIFT(1); break;
      case 2: HEL_SHIFT(2); break;
      default: HEL_SHIFT(3); break;
    }
  } else {
    switch (wv) {
      case 0: HEL_SHIFT(0); break;
      case 1: HEL_SHIFT(1); break;
      case 2: HEL_SHIFT(2); break;
      case 3: HEL_SHIFT(3); break;
      case 4: HEL_SHIFT(4); break;
      case 5: HEL_SHIFT(5); break;
      case 6: HEL_SHIFT(6); break;
      default: HEL_SHIFT(7); break;
    }
  }
#undef HEL_SHIFT
}

template <typename T, int N>
static int launch_hel_shift(int mode, const uint32_t *parents, unsigned npar,
                            const uint32_t *cstart, const uint32_t *cend, const uint32_t *cflag,
                            const T *in, T *out, const void *plan, const T *A,
                            const T *A2, const T *CS, const T *zero,
                            const wbfmm_m2l_plan_t *hplan) {
  if (!plan_matches<N>(hplan)) {
    fprintf(stderr, "wbfmm (HIP): shift plan for order %d does not match the kernel's bins\n", N);
    return WBFMM_HIP_ERR_ARG;
  }
  constexpr int W = N >= 16 ? 8 : 4;
  const dim3 grid((npar + 7) / 8), block(64 * W);
  if (mode == 0)
    WBFMM_LAUNCH(sizeof(T) == 4 ? "hel_m2m_f32" : "hel_m2m", (k_hel_shift_mfma<T, N, W, 0>), grid,
                 block, 0, parents, npar, cstart,
                 cend, cflag, (const T2t<T> *)in, (T2t<T> *)out,
                 (const wbfmm_m2l_plan_t *)plan, A, A2, CS, (const T2t<T> *)zero);
  else
    WBFMM_LAUNCH(sizeof(T) == 4 ? "hel_l2l_f32" : "hel_l2l", (k_hel_shift_mfma<T, N, W, 1>), grid,
                 block, 0, parents, npar, cstart,
                 cend, cflag, (const T2t<T> *)in, (T2t<T> *)out,
                 (const wbfmm_m2l_plan_t *)plan, A, A2, CS, (const T2t<T> *)zero);
  return 0;
}


template <typename T, int n, int HI>
int hel_m2l_orders_t(int N, const uint32_t *cls, const ClassArgs &ca, unsigned nch,
                     const uint32_t *occ, int level, const T *mps, T *mpr, const void *plan,
                     const T *A, const T *A2, const T *CS, const T *zero,
                     const wbfmm_m2l_plan_t *hplan) {
  if constexpr (n > HI) {
    return WBFMM_HIP_ERR_ORDER;
  } else {
    if (N == n)
      return launch_hel_m2l<T, n>(cls, ca, nch, occ, level, mps, mpr, plan, A, A2, CS, zero,
                                  hplan);
    return hel_m2l_orders_t<T, n + 1, HI>(N, cls, ca, nch, occ, level, mps, mpr, plan, A, A2,
                                          CS, zero, hplan);
  }
}

template <typename T, int n, int HI>
int hel_shift_orders_t(int mode, int N, const uint32_t *parents, unsigned npar,
                       const uint32_t *cstart, const uint32_t *cend, const uint32_t *cflag,
                       const T *in, T *out, const void *plan, const T *A, const T *A2,
                       const T *CS, const T *zero, const wbfmm_m2l_plan_t *hplan) {
  if constexpr (n > HI) {
    return WBFMM_HIP_ERR_ORDER;
  } else {
    if (N == n)
      return launch_hel_shift<T, n>(mode, parents, npar, cstart, cend, cflag, in, out, plan, A,
                                    A2, CS, zero, hplan);
    return hel_shift_orders_t<T, n + 1, HI>(mode, N, parents, npar, cstart, cend, cflag, in,
                                            out, plan, A, A2, CS, zero, hplan);
  }
}

#define WBFMM_HEL_RANGE_DEFINE_T(LO, HI, T)                                                  \
  int hel_m2l_range_##LO##_##HI(int N, const uint32_t *cls, const ClassArgs &ca,            \
                                unsigned nch, const uint32_t *occ, int level, const T *mps, \
                                T *mpr, const void *plan, const T *A, const T *A2,          \
                                const T *CS, const T *zero, const wbfmm_m2l_plan_t *hplan) { \
    return hel_m2l_orders_t<T, LO, HI>(N, cls, ca, nch, occ, level, mps, mpr, plan, A, A2,  \
                                       CS, zero, hplan);                                    \
  }                                                                                         \
  int hel_shift_range_##LO##_##HI(int mode, int N, const uint32_t *parents, unsigned npar,  \
                                  const uint32_t *cstart, const uint32_t *cend,             \
                                  const uint32_t *cflag, const T *in, T *out,               \
                                  const void *plan, const T *A, const T *A2, const T *CS,   \
                                  const T *zero, const wbfmm_m2l_plan_t *hplan) {           \
    return hel_shift_orders_t<T, LO, HI>(mode, N, parents, npar, cstart, cend, cflag, in,   \
                                         out, plan, A, A2, CS, zero, hplan);                \
  }
#define WBFMM_HEL_RANGE_DEFINE(LO, HI)       \
  WBFMM_HEL_RANGE_DEFINE_T(LO, HI, double)   \
  WBFMM_HEL_RANGE_DEFINE_T(LO, HI, float)

}  // namespace wbfmm

#endif
