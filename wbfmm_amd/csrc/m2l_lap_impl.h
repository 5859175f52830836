// m2l_lap_impl.h -- the Laplace matrix-core translation kernels (M2L,
// M2M, L2L) and their launchers, as templates over the precision and the
// expansion order.  Each m2l_lap_o*.hip translation unit instantiates one
// range of orders (m2l_lap_ranges.h), so the orders compile in parallel;
// m2l_mfma.hip holds the entry points and dispatches by order.
//
// Design of the M2L form: one wavefront (or a block of W sharing the
// work) takes 16 target boxes with the same Morton child digit and loops
// over the list-4 offsets; see m2l_mfma.hip for the plan and the operand
// layout.

#ifndef WBFMM_M2L_LAP_IMPL_H
#define WBFMM_M2L_LAP_IMPL_H

#include <cstdio>
#include <cstdlib>
#include <cstring>

#include <algorithm>

#include <hip/hip_runtime.h>

#include "hip_common.h"
#include "m2l_mfma.h"
#include "m2l_lap_ranges.h"

namespace wbfmm {

template <typename T, int N>
__global__ __launch_bounds__(64) void k_lap_m2l_mfma(
    const uint32_t *__restrict__ cls, ClassArgs ca,
    const uint32_t *__restrict__ occ, int level,
    const T *__restrict__ mps, T *__restrict__ mpr,
    const wbfmm_m2l_plan_t *__restrict__ P, const T *__restrict__ A1,
    const T *__restrict__ A3, const T *__restrict__ A2,
    const T *__restrict__ CS, const T *__restrict__ zero, int nsplit,
    T *__restrict__ part, unsigned ntot) {
  typedef typename MF<T>::t2 T2;
  typedef typename MF<T>::v4 V4;
  using SH = M2LShape<N>;
  constexpr Bins BN = SH::BN;
  constexpr int NB = SH::NB, KT = SH::KT, C = SH::C, NU = SH::NU, S = SH::S;
  constexpr int NAU = SH::NAU;
  __shared__ T R[(S + 1) * 16];  // slot S: dummy destination
  // slot maps transposed so that one lane reads its four rows (q = 0..3)
  // of a bin at once: tmapX[4 b + grp] = slots of rows 16 b + drow(grp, q)
  __shared__ int4 tmap1[4 * NB], tmap2[4 * NB], tmap3[4 * NB];
  __shared__ int4 uinf[NU];  // packed position of re, m, stage-1 re slot, im slot
  __shared__ T2 cs[2][N + 1];  // cos, sin of m chi: current / next offset
  __shared__ unsigned vm[16][7];
  __shared__ unsigned anym[7];

  const int lane = threadIdx.x, col = lane & 15, grp = lane >> 4;

  for (int i = lane; i < 4 * NB; i += 64) {
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
  for (int u = lane; u < NU; u += 64) {
    int n = 0;
    while ((n + 1) * (n + 2) / 2 <= u) n++;
    const int m = u - n * (n + 1) / 2;
    uinf[u] = make_int4(n * n + (m ? 2 * m - 1 : 0), m, P->s1[u] & 0xffff,
                        (P->s1[u] >> 16) & 0xffff);
  }
  for (int i = lane; i < (S + 1) * 16; i += 64) R[i] = T(0);
  for (int i = lane; i < 16 * 7; i += 64) vm[i / 7][i % 7] = 0u;

  // block -> chunk of 16 targets; consecutive chunks stay on one XCD
  // nsplit > 1 (levels with few targets): nsplit blocks share a chunk,
  // block `split` taking the cube indices k = split (mod nsplit), and write
  // partial sums that k_m2l_reduce adds in a fixed order
  const int split = (int)(blockIdx.x % (unsigned)nsplit);
  int c;
  unsigned cj;
  // a block with no chunk leaves before any barrier (uniform over the block)
  if (!block_chunk(blockIdx.x / (unsigned)nsplit, ca, c, cj)) return;
  const unsigned k16 = cj * 16 + col;
  const bool tvalid = k16 < ca.count[c];
  const uint32_t tb = tvalid ? cls[ca.start[c] + k16] : 0u;
  const int ti = compact3(tb), tj = compact3(tb >> 1), tk = compact3(tb >> 2);
  const int px = c & 1, py = (c >> 1) & 1, pz = (c >> 2) & 1;
  const int nbox = 1 << level;
  __syncthreads();

  // list-4 sources of each target, as bits over the 6x6x6 cube of the
  // parent's neighbours' children, k = 36 (dz+2+pz) + 6 (dy+2+py) + dx+2+px.
  // dx varies fastest: the 16 targets of a wave are consecutive in Morton
  // order, i.e. spread along x, so offsets two steps apart in dx read
  // mostly the same source boxes while they are still in L2.
  for (int k = grp; k < 216; k += 4) {
    const int dx = k % 6 - 2 - px, dy = (k / 6) % 6 - 2 - py, dz = k / 36 - 2 - pz;
    if (dx >= -1 && dx <= 1 && dy >= -1 && dy <= 1 && dz >= -1 && dz <= 1) continue;
    const int si = ti + dx, sj = tj + dy, sk = tk + dz;
    if (!tvalid || si < 0 || si >= nbox || sj < 0 || sj >= nbox || sk < 0 || sk >= nbox)
      continue;
    const uint64_t sb = spread3(si) | (spread3(sj) << 1) | (spread3(sk) << 2);
    if ((occ[sb >> 5] >> (sb & 31)) & 1u) atomicOr(&vm[col][k >> 5], 1u << (k & 31));
  }
  __syncthreads();
  if (lane < 7) {
    unsigned a = 0;
    for (int j = 0; j < 16; j++) a |= vm[j][lane];
    anym[lane] = a;
  }
  __syncthreads();

  auto next_k = [&](int k) {
    for (k++; k < 216; k++)
      if (((anym[k >> 5] >> (k & 31)) & 1u) && k % nsplit == split) break;
    return k;
  };
  auto offset_of = [&](int k) {
    const int dx = k % 6 - 2 - px, dy = (k / 6) % 6 - 2 - py, dz = k / 36 - 2 - pz;
    return (dx + 3) * 49 + (dy + 3) * 7 + dz + 3;
  };
  // this lane's source expansion for cube index k (zeros when empty) and
  // the offset's azimuthal factors into cs[p]
  auto source_of = [&](int k, int p) {
    const int dx = k % 6 - 2 - px, dy = (k / 6) % 6 - 2 - py, dz = k / 36 - 2 - pz;
    if (lane <= N) {
      const T *t = CS + ((size_t)offset_of(k) * (N + 1) + lane) * 2;
      cs[p][lane] = mk2<T>(t[0], t[1]);
    }
    const bool v = (vm[col][k >> 5] >> (k & 31)) & 1u;
    const uint32_t sb = morton_add(tb, dilate_offset(dx, 0, kMortonX),
                                   dilate_offset(dy, 1, kMortonY), dilate_offset(dz, 2, kMortonZ));
    return v ? mps + (uint64_t)sb * C : zero;
  };
  // Units u = grp + 4 i of column col: the lane loads, rotates (Zin) and
  // accumulates the same units.  The last lanes repeat unit NU-1 (same
  // value to the same slot) and are not written out.  The im part is
  // always loaded (for m = 0 it is the next coefficient, multiplied by
  // sin 0 = 0 and written to the dummy slot).
  T pre_re[NAU], pre_im[NAU];
  // units [i0, i1) of this lane
  auto load_sources = [&](const T *src, const int4 *tu, int gp, int i0, int i1) {
#pragma unroll
    for (int i = i0; i < i1; i++) {
      const int u = min(gp + 4 * i, NU - 1);
      const T *q = src + tu[u].x;
      pre_re[i] = q[0];
      pre_im[i] = q[1];
    }
  };
  auto deposit = [&](const int4 *tu, int gp, int cl, int p, int i0, int i1) {
#pragma unroll
    for (int i = i0; i < i1; i++) {
      const int u = min(gp + 4 * i, NU - 1);
      const int4 inf = tu[u];
      const T2 e = cs[p][inf.y];
      R[inf.z * 16 + cl] = e.x * pre_re[i] - e.y * pre_im[i];
      R[inf.w * 16 + cl] = e.y * pre_re[i] + e.x * pre_im[i];
    }
  };
  // without the register prefetch: load and deposit in groups of 8 units
  auto load_deposit = [&](const T *src, const int4 *tu, int gp, int cl, int p) {
#pragma unroll
    for (int i0 = 0; i0 < NAU; i0 += 8) {
      const int i1 = i0 + 8 < NAU ? i0 + 8 : NAU;
      load_sources(src, tu, gp, i0, i1);
      deposit(tu, gp, cl, p, i0, i1);
    }
  };

  T acc_re[NAU], acc_im[NAU];
#pragma unroll
  for (int i = 0; i < NAU; i++) acc_re[i] = acc_im[i] = T(0);

  int k = next_k(-1), p = 0;
  if (k < 216) load_deposit(source_of(k, 0), uinf, grp, col, 0);
  while (k < 216) {
    // Views of the constant LDS tables and lane indices through an offset
    // the compiler cannot see (0, from inline asm, fresh every iteration):
    // the per-lane table values are then re-read where they are used
    // instead of being hoisted out of the loop into registers.
    int z0;
    asm volatile("s_mov_b32 %0, 0" : "=s"(z0));
    const int4 *t1 = tmap1 + z0, *t2 = tmap2 + z0, *t3 = tmap3 + z0, *tu = uinf + z0;
    const int ln = lane + z0, cl = ln & 15, gp = ln >> 4;
    const int o = offset_of(k);
    const int th = __builtin_amdgcn_readfirstlane(P->th[o]);
    const int ri = __builtin_amdgcn_readfirstlane(P->rix[o]);
    const int kn = next_k(k);

    // stage 1: rotate into the offset's frame
    V4 y[NB];
    stage<T, NB, KT>(BN, A1 + (size_t)th * KT * 64, R, ln, gp, cl, y);
#pragma unroll
    for (int b = 0; b < NB; b++) {
      const int4 d = t1[4 * b + gp];
      R[d.x * 16 + cl] = y[b][0];
      R[d.y * 16 + cl] = y[b][1];
      R[d.z * 16 + cl] = y[b][2];
      R[d.w * 16 + cl] = y[b][3];
    }
    if (SH::PREFETCH && kn < 216) load_sources(source_of(kn, p ^ 1), tu, gp, 0, NAU);

    // stage 2: coaxial S|R at this offset's distance
    stage<T, NB, KT>(BN, A2 + (size_t)ri * KT * 64, R, ln, gp, cl, y);
#pragma unroll
    for (int b = 0; b < NB; b++) {
      const int4 d = t2[4 * b + gp];
      R[d.x * 16 + cl] = y[b][0];
      R[d.y * 16 + cl] = y[b][1];
      R[d.z * 16 + cl] = y[b][2];
      R[d.w * 16 + cl] = y[b][3];
    }

    // stage 3: rotate back (without the azimuthal factor)
    stage<T, NB, KT>(BN, A3 + (size_t)th * KT * 64, R, ln, gp, cl, y);
#pragma unroll
    for (int b = 0; b < NB; b++) {
      const int4 d = t3[4 * b + gp];
      R[d.x * 16 + cl] = y[b][0];
      R[d.y * 16 + cl] = y[b][1];
      R[d.z * 16 + cl] = y[b][2];
      R[d.w * 16 + cl] = y[b][3];
    }

    // Zout and accumulate
#pragma unroll
    for (int i = 0; i < NAU; i++) {
      const int u = min(gp + 4 * i, NU - 1);
      const T2 e = cs[p][tu[u].y];
      const T tr = R[u * 16 + cl], tq = R[(NU + u) * 16 + cl];
      acc_re[i] += e.x * tr + e.y * tq;
      acc_im[i] += e.x * tq - e.y * tr;
    }

    if (kn < 216) {
      if (SH::PREFETCH)
        deposit(tu, gp, cl, p ^ 1, 0, NAU);
      else
        load_deposit(source_of(kn, p ^ 1), tu, gp, cl, p ^ 1);
    }
    p ^= 1;
    k = kn;
  }

  if (!tvalid) return;
  if (nsplit > 1) {
    // every coefficient of the partial is written (m = 0 has no im part)
    T *dst = part + ((size_t)split * ntot + ca.start[c] + k16) * C;
#pragma unroll
    for (int i = 0; i < NAU; i++) {
      const int u = grp + 4 * i;
      if (u < NU) {
        const int4 inf = uinf[u];
        dst[inf.x] = acc_re[i];
        if (inf.y) dst[inf.x + 1] = acc_im[i];
      }
    }
    return;
  }
  T *out = mpr + (size_t)tb * C;
#pragma unroll
  for (int i = 0; i < NAU; i++) {
    const int u = grp + 4 * i;
    if (u < NU) {
      const int4 inf = uinf[u];
      out[inf.x] += acc_re[i];
      if (inf.y) out[inf.x + 1] += acc_im[i];
    }
  }
}

// ---------------------------------------------------------------------
// W wavefronts per chunk (W = 2 or 4).  The one-wavefront kernel holds ~490
// registers at order 12 (stage outputs, the next offset's sources and the
// 16 local expansions), so one wavefront runs per SIMD and nothing hides
// its LDS and memory waits.  Here a block of W wavefronts shares one LDS
// buffer R for its 16 targets: wavefront w computes the bins b = w (mod W)
// of every stage and owns the units u = grp + 4 w + 4 W i for loading,
// rotation (Zin) and accumulation, so each holds about 1/W of the registers
// and several wavefronts run per SIMD.  Barriers separate the reads of R by
// a stage from the writes of its results.  Every output takes the same
// products in the same order as in the one-wavefront kernel; the kernels
// may differ by the rounding of contracted (fused) expressions, and each
// is deterministic: every LDS slot has one writer.
// ---------------------------------------------------------------------

template <typename T, int N, int W, int WI>
__device__ __forceinline__ void m2l_w2_loop(
    T *R, const int4 *tmap1, const int4 *tmap2, const int4 *tmap3, const int4 *uinf,
    typename MF<T>::t2 (*cs)[N + 1], unsigned (*vm)[7], const unsigned *anym, int lane,
    int col, int grp, bool tvalid, uint32_t tb, int ti, int tj, int tk, int px, int py,
    int pz, int c, unsigned k16, int split, int nsplit, const ClassArgs &ca,
    const wbfmm_m2l_plan_t *__restrict__ P, const T *__restrict__ A1,
    const T *__restrict__ A3, const T *__restrict__ A2, const T *__restrict__ CS,
    const T *__restrict__ zero, const T *__restrict__ mps, T *__restrict__ mpr,
    T *__restrict__ part, unsigned ntot) {
  typedef typename MF<T>::t2 T2;
  typedef typename MF<T>::v4 V4;
  using SH = M2LShape<N>;
  constexpr int C = SH::C, NU = SH::NU;
  constexpr bool DB = M2LDouble<T, N>::value;
  T *R1 = R + (SH::S + 1) * 16;  // the second buffer (DB only)
  constexpr int NAW = (NU + 4 * W - 1) / (4 * W);  // units per lane
  constexpr WaveSeq WS = make_seq(SH::BN, W, WI);
  constexpr int NBW = WS.nb;
  // register prefetch of the next offset's sources while two wavefronts
  // per SIMD still fit in the register file
  constexpr bool PREFETCH = true;
  const int ubase = grp + 4 * WI;

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
  // class order (ca.ord): the position after pos in kOffOrder[c] whose
  // offset is used (189 when there is none)
  auto next_pos = [&](int pos) {
    for (pos++; pos < 189; pos++) {
      const int k = kOffOrder[c].k[pos];
      if (((anym[k >> 5] >> (k & 31)) & 1u) && (nsplit == 1 || k % nsplit == split)) break;
    }
    return pos;
  };
  const bool cord = ca.ord != 0;
  auto offset_of = [&](int k) {
    const int dx = k % 6 - 2 - px, dy = (k / 6) % 6 - 2 - py, dz = k / 36 - 2 - pz;
    return (dx + 3) * 49 + (dy + 3) * 7 + dz + 3;
  };
  // both wavefronts write the offset's factors (the same values)
  auto source_of = [&](int k, int p) {
    const int dx = k % 6 - 2 - px, dy = (k / 6) % 6 - 2 - py, dz = k / 36 - 2 - pz;
    if (lane <= N) {
      const T *t = CS + ((size_t)offset_of(k) * (N + 1) + lane) * 2;
      cs[p][lane] = mk2<T>(t[0], t[1]);
    }
    const bool v = (vm[col][k >> 5] >> (k & 31)) & 1u;
    const uint32_t sb = morton_add(tb, dilate_offset(dx, 0, kMortonX),
                                   dilate_offset(dy, 1, kMortonY), dilate_offset(dz, 2, kMortonZ));
    return v ? mps + (uint64_t)sb * C : zero;
  };
  T pre_re[NAW], pre_im[NAW];
  auto load_sources = [&](const T *src, const int4 *tu, int ub, int i0, int i1) {
#pragma unroll
    for (int i = i0; i < i1; i++) {
      const int u = min(ub + 4 * W * i, NU - 1);
      const T *q = src + tu[u].x;
      pre_re[i] = q[0];
      pre_im[i] = q[1];
    }
  };
  // Lanes past the last unit (loaded as unit NU-1) write nothing: the two
  // wavefronts run separately compiled code, whose roundings of the same
  // expression may differ, so a slot must have a single writer.
  auto deposit = [&](const int4 *tu, int ub, int cl, int p, int i0, int i1) {
#pragma unroll
    for (int i = i0; i < i1; i++) {
      if (ub + 4 * W * i >= NU) continue;
      const int4 inf = tu[ub + 4 * W * i];
      const T2 e = cs[p][inf.y];
      R[inf.z * 16 + cl] = e.x * pre_re[i] - e.y * pre_im[i];
      R[inf.w * 16 + cl] = e.y * pre_re[i] + e.x * pre_im[i];
    }
  };
  auto load_deposit = [&](const T *src, const int4 *tu, int ub, int cl, int p) {
#pragma unroll
    for (int i0 = 0; i0 < NAW; i0 += 8) {
      const int i1 = i0 + 8 < NAW ? i0 + 8 : NAW;
      load_sources(src, tu, ub, i0, i1);
      deposit(tu, ub, cl, p, i0, i1);
    }
  };
  auto write_out = [&](const int4 *tm, const V4 *y, int gp, int cl) {
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
  int pos = cord ? __builtin_amdgcn_readfirstlane(next_pos(-1)) : 0;
  int k = cord ? (pos < 189 ? (int)kOffOrder[c].k[pos] : 216)
               : __builtin_amdgcn_readfirstlane(next_k(-1));
  int p = 0;
  int th = 0, ri = 0;
  if (k < 216) {
    th = __builtin_amdgcn_readfirstlane(P->th[offset_of(k)]);
    ri = __builtin_amdgcn_readfirstlane(P->rix[offset_of(k)]);
    load_deposit(source_of(k, 0), uinf, ubase, col, 0);
  }
  // stage operands held across stages (StageOps) where they fit
  constexpr bool PF = StageOps<N, W, WI>::fits;
  T av[StageOps<N, W, WI>::n];
  if constexpr (PF) ops_load<T, N, W, WI>(av, A1 + (size_t)th * SH::KT * 64, lane);
  __syncthreads();
  while (k < 216) {
    int z0;
    asm volatile("s_mov_b32 %0, 0" : "=s"(z0));
    const int4 *t1 = tmap1 + z0, *t2 = tmap2 + z0, *t3 = tmap3 + z0, *tu = uinf + z0;
    const int ln = lane + z0, cl = ln & 15, gp = ln >> 4, ub = gp + 4 * WI;
    int pos_n = 0, kn;
    if (cord) {
      pos_n = __builtin_amdgcn_readfirstlane(next_pos(pos));
      kn = pos_n < 189 ? (int)kOffOrder[c].k[pos_n] : 216;
    } else {
      kn = __builtin_amdgcn_readfirstlane(next_k(k));
    }
    int th_n = 0, ri_n = 0;
    if (kn < 216) {
      th_n = __builtin_amdgcn_readfirstlane(P->th[offset_of(kn)]);
      ri_n = __builtin_amdgcn_readfirstlane(P->rix[offset_of(kn)]);
      // the next offset's sources are requested before this offset's
      // stages: the registers were emptied by the last deposit, and
      // cs[p ^ 1] was last read before the barrier that ended the last offset
      if (PREFETCH) load_sources(source_of(kn, p ^ 1), tu, ub, 0, NAW);
    }

    // double-buffered: R -> R1 -> R -> R1, one barrier per stage; the
    // deposit of the next offset goes into R while R1 is read below
    T *RZ = R;
    if constexpr (DB && PF) {
      stage_wop<T, N, W, WI>(av, A2 + (size_t)ri * SH::KT * 64, R, R1, t1, ln, gp, cl);
      __syncthreads();
      stage_wop<T, N, W, WI>(av, A3 + (size_t)th * SH::KT * 64, R1, R, t2, ln, gp, cl);
      __syncthreads();
      stage_wop<T, N, W, WI>(av, A1 + (size_t)th_n * SH::KT * 64, R, R1, t3, ln, gp, cl);
      __syncthreads();
      RZ = R1;
    } else if constexpr (PF) {
      V4 y[NBW > 0 ? NBW : 1];
      stage_wp<T, N, W, WI>(av, A2 + (size_t)ri * SH::KT * 64, R, ln, gp, cl, y);
      __syncthreads();
      write_out(t1, y, gp, cl);
      __syncthreads();
      stage_wp<T, N, W, WI>(av, A3 + (size_t)th * SH::KT * 64, R, ln, gp, cl, y);
      __syncthreads();
      write_out(t2, y, gp, cl);
      __syncthreads();
      stage_wp<T, N, W, WI>(av, A1 + (size_t)th_n * SH::KT * 64, R, ln, gp, cl, y);
      __syncthreads();
      write_out(t3, y, gp, cl);
      __syncthreads();
    } else if constexpr (DB) {
      stage_wo<T, N, W, WI>(A1 + (size_t)th * SH::KT * 64, R, R1, t1, ln, gp, cl);
      __syncthreads();
      stage_wo<T, N, W, WI>(A2 + (size_t)ri * SH::KT * 64, R1, R, t2, ln, gp, cl);
      __syncthreads();
      stage_wo<T, N, W, WI>(A3 + (size_t)th * SH::KT * 64, R, R1, t3, ln, gp, cl);
      __syncthreads();
      RZ = R1;
    } else {
      V4 y[NBW > 0 ? NBW : 1];
      stage_w<T, N, W, WI>(A1 + (size_t)th * SH::KT * 64, R, ln, gp, cl, y);
      __syncthreads();
      write_out(t1, y, gp, cl);
      __syncthreads();

      stage_w<T, N, W, WI>(A2 + (size_t)ri * SH::KT * 64, R, ln, gp, cl, y);
      __syncthreads();
      write_out(t2, y, gp, cl);
      __syncthreads();

      stage_w<T, N, W, WI>(A3 + (size_t)th * SH::KT * 64, R, ln, gp, cl, y);
      __syncthreads();
      write_out(t3, y, gp, cl);
      __syncthreads();
    }

#pragma unroll
    for (int i = 0; i < NAW; i++) {
      const int u = min(ub + 4 * W * i, NU - 1);
      const T2 e = cs[p][tu[u].y];
      const T tr = RZ[u * 16 + cl], tq = RZ[(NU + u) * 16 + cl];
      acc_re[i] += e.x * tr + e.y * tq;
      acc_im[i] += e.x * tq - e.y * tr;
    }
    if constexpr (!DB) __syncthreads();  // every read of R done before the next deposit

    if (kn < 216) {
      if (PREFETCH)
        deposit(tu, ub, cl, p ^ 1, 0, NAW);
      else
        load_deposit(source_of(kn, p ^ 1), tu, ub, cl, p ^ 1);
    }
    __syncthreads();
    p ^= 1;
    k = kn;
    pos = pos_n;
    th = th_n;
    ri = ri_n;
  }

  if (!tvalid) return;
  if (nsplit > 1) {
    T *dst = part + ((size_t)split * ntot + ca.start[c] + k16) * C;
#pragma unroll
    for (int i = 0; i < NAW; i++) {
      const int u = ubase + 4 * W * i;
      if (u < NU) {
        const int4 inf = uinf[u];
        dst[inf.x] = acc_re[i];
        if (inf.y) dst[inf.x + 1] = acc_im[i];
      }
    }
    return;
  }
  T *out = mpr + (size_t)tb * C;
#pragma unroll
  for (int i = 0; i < NAW; i++) {
    const int u = ubase + 4 * W * i;
    if (u < NU) {
      const int4 inf = uinf[u];
      out[inf.x] += acc_re[i];
      if (inf.y) out[inf.x + 1] += acc_im[i];
    }
  }
}

// Wavefronts per SIMD the register file must hold: at FP64 orders 9..12
// with eight wavefronts, three blocks per CU (the double LDS buffer of
// about 50 KB leaves room for three, and 80 registers per lane give six
// wavefronts per SIMD); otherwise two where the LDS buffer leaves room for
// them (FP32, and FP64 up to order 15); above, LDS limits the blocks per
// CU anyway and the registers are not capped
template <typename T, int N, int W>
constexpr int m2l_waves_per_eu() {
  return (sizeof(T) == 8 && W == 8 && N >= 9 && N <= 12) ? 6
         : (sizeof(T) == 4 || N <= 15)                   ? 2
                                                         : 1;
}

template <typename T, int N, int W>
__global__ __launch_bounds__(64 * W) __attribute__((amdgpu_waves_per_eu(
    m2l_waves_per_eu<T, N, W>()))) void k_lap_m2l_mfma2(
    const uint32_t *__restrict__ cls, ClassArgs ca,
    const uint32_t *__restrict__ occ, int level,
    const T *__restrict__ mps, T *__restrict__ mpr,
    const wbfmm_m2l_plan_t *__restrict__ P, const T *__restrict__ A1,
    const T *__restrict__ A3, const T *__restrict__ A2,
    const T *__restrict__ CS, const T *__restrict__ zero, int nsplit,
    T *__restrict__ part, unsigned ntot) {
  typedef typename MF<T>::t2 T2;
  using SH = M2LShape<N>;
  constexpr int NB = SH::NB, NU = SH::NU, S = SH::S;
  __shared__ T R[(M2LDouble<T, N>::value ? 2 : 1) * (S + 1) * 16];
  __shared__ int4 tmap1[4 * NB], tmap2[4 * NB], tmap3[4 * NB];
  __shared__ int4 uinf[NU];
  __shared__ T2 cs[2][N + 1];
  __shared__ unsigned vm[16][7];
  __shared__ unsigned anym[7];

  // wv through readfirstlane: a scalar value, so the branch into the
  // per-wavefront code below is a scalar branch and each wavefront runs
  // (and meets the barriers of) its own path only
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
    uinf[u] = make_int4(n * n + (m ? 2 * m - 1 : 0), m, P->s1[u] & 0xffff,
                        (P->s1[u] >> 16) & 0xffff);
  }
  for (int i = tid; i < (M2LDouble<T, N>::value ? 2 : 1) * (S + 1) * 16; i += 64 * W) R[i] = T(0);
  for (int i = tid; i < 16 * 7; i += 64 * W) vm[i / 7][i % 7] = 0u;

  const int split = (int)(blockIdx.x % (unsigned)nsplit);
  int c;
  unsigned cj;
  // a block with no chunk leaves before any barrier (uniform over the block)
  if (!block_chunk(blockIdx.x / (unsigned)nsplit, ca, c, cj)) return;
  const unsigned k16 = cj * 16 + col;
  const bool tvalid = k16 < ca.count[c];
  const uint32_t tb = tvalid ? cls[ca.start[c] + k16] : 0u;
  const int ti = compact3(tb), tj = compact3(tb >> 1), tk = compact3(tb >> 2);
  const int px = c & 1, py = (c >> 1) & 1, pz = (c >> 2) & 1;
  const int nbox = 1 << level;
  __syncthreads();

  for (int k = grp + 4 * wv; k < 216; k += 4 * W) {
    const int dx = k % 6 - 2 - px, dy = (k / 6) % 6 - 2 - py, dz = k / 36 - 2 - pz;
    if (dx >= -1 && dx <= 1 && dy >= -1 && dy <= 1 && dz >= -1 && dz <= 1) continue;
    const int si = ti + dx, sj = tj + dy, sk = tk + dz;
    if (!tvalid || si < 0 || si >= nbox || sj < 0 || sj >= nbox || sk < 0 || sk >= nbox)
      continue;
    const uint64_t sb = spread3(si) | (spread3(sj) << 1) | (spread3(sk) << 2);
    if ((occ[sb >> 5] >> (sb & 31)) & 1u) atomicOr(&vm[col][k >> 5], 1u << (k & 31));
  }
  __syncthreads();
  if (tid < 7) {
    unsigned a = 0;
    for (int j = 0; j < 16; j++) a |= vm[j][tid];
    anym[tid] = a;
  }
  __syncthreads();

  if constexpr (W == 2) {
    if (wv == 0)
      m2l_w2_loop<T, N, 2, 0>(R, tmap1, tmap2, tmap3, uinf, cs, vm, anym, lane, col, grp, tvalid, tb,
                            ti, tj, tk, px, py, pz, c, k16, split, nsplit, ca, P, A1, A3, A2,
                            CS, zero, mps, mpr, part, ntot);
    else
      m2l_w2_loop<T, N, 2, 1>(R, tmap1, tmap2, tmap3, uinf, cs, vm, anym, lane, col, grp, tvalid, tb,
                            ti, tj, tk, px, py, pz, c, k16, split, nsplit, ca, P, A1, A3, A2,
                            CS, zero, mps, mpr, part, ntot);
  } else if constexpr (W == 8) {
#define CHUNK8(WI)                                                                            \
  m2l_w2_loop<T, N, 8, WI>(R, tmap1, tmap2, tmap3, uinf, cs, vm, anym, lane, col, grp, tvalid, \
                           tb, ti, tj, tk, px, py, pz, c, k16, split, nsplit, ca, P, A1, A3,  \
                           A2, CS, zero, mps, mpr, part, ntot)
    switch (wv) {
      case 0: CHUNK8(0); break;
      case 1: CHUNK8(1); break;
      case 2: CHUNK8(2); break;
      case 3: CHUNK8(3); break;
      case 4: CHUNK8(4); break;
      case 5: CHUNK8(5); break;
      case 6: CHUNK8(6); break;
      default: CHUNK8(7); break;
    }
#undef CHUNK8
  } else {
    switch (wv) {
      case 0:
        m2l_w2_loop<T, N, W, 0>(R, tmap1, tmap2, tmap3, uinf, cs, vm, anym, lane, col, grp, tvalid, tb,
                            ti, tj, tk, px, py, pz, c, k16, split, nsplit, ca, P, A1, A3, A2,
                            CS, zero, mps, mpr, part, ntot);
        break;
      case 1:
        m2l_w2_loop<T, N, W, 1>(R, tmap1, tmap2, tmap3, uinf, cs, vm, anym, lane, col, grp, tvalid, tb,
                            ti, tj, tk, px, py, pz, c, k16, split, nsplit, ca, P, A1, A3, A2,
                            CS, zero, mps, mpr, part, ntot);
        break;
      case 2:
        m2l_w2_loop<T, N, W, 2>(R, tmap1, tmap2, tmap3, uinf, cs, vm, anym, lane, col, grp, tvalid, tb,
                            ti, tj, tk, px, py, pz, c, k16, split, nsplit, ca, P, A1, A3, A2,
                            CS, zero, mps, mpr, part, ntot);
        break;
      default:
        m2l_w2_loop<T, N, W, 3>(R, tmap1, tmap2, tmap3, uinf, cs, vm, anym, lane, col, grp, tvalid, tb,
                            ti, tj, tk, px, py, pz, c, k16, split, nsplit, ca, P, A1, A3, A2,
                            CS, zero, mps, mpr, part, ntot);
        break;
    }
  }
}


// ---------------------------------------------------------------------
// M2M (MODE 0, src/shift-laplace.c:225-442) and L2L (MODE 1, :562-789) in
// the same split MFMA form: a wavefront takes 16 parents and, for each of
// the 8 child directions, translates the 16 vectors at once.  M2M sums
// the children of each parent in registers and writes the parent; L2L
// translates each parent's local expansion to child c and adds it there.
// Child direction c uses polar-angle table th(c) (0: acos(1/sqrt3), 1:
// acos(-1/sqrt3)) and the azimuthal factors CS[c].
// ---------------------------------------------------------------------
template <typename T, int N, int MODE>
__global__ __launch_bounds__(64) void k_lap_shift_mfma(
    const uint32_t *__restrict__ parents, unsigned npar,
    const uint32_t *__restrict__ cstart, const uint32_t *__restrict__ cend,
    const uint32_t *__restrict__ cflag, const T *__restrict__ in,
    T *__restrict__ out, const wbfmm_m2l_plan_t *__restrict__ P,
    const T *__restrict__ A1, const T *__restrict__ A3,
    const T *__restrict__ A2, const T *__restrict__ CS,
    const T *__restrict__ zero) {
  typedef typename MF<T>::t2 T2;
  typedef typename MF<T>::v4 V4;
  using SH = M2LShape<N>;
  constexpr Bins BN = SH::BN;
  constexpr int NB = SH::NB, KT = SH::KT, C = SH::C, NU = SH::NU, S = SH::S;
  constexpr int NAU = SH::NAU;
  __shared__ T R[(S + 1) * 16];
  __shared__ int4 tmap1[4 * NB], tmap2[4 * NB], tmap3[4 * NB];
  __shared__ int4 uinf[NU];
  __shared__ T2 cs[N + 1];

  const int lane = threadIdx.x, col = lane & 15, grp = lane >> 4;
  for (int i = lane; i < 4 * NB; i += 64) {
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
  for (int u = lane; u < NU; u += 64) {
    int n = 0;
    while ((n + 1) * (n + 2) / 2 <= u) n++;
    const int m = u - n * (n + 1) / 2;
    uinf[u] = make_int4(n * n + (m ? 2 * m - 1 : 0), m, P->s1[u] & 0xffff,
                        (P->s1[u] >> 16) & 0xffff);
  }
  for (int i = lane; i < (S + 1) * 16; i += 64) R[i] = T(0);
  const unsigned pi = blockIdx.x * 16 + col;
  const bool pvalid = pi < npar;
  const uint64_t p = pvalid ? parents[pi] : 0;
  __syncthreads();

  T pre_re[NAU], pre_im[NAU], acc_re[NAU], acc_im[NAU];
  auto load = [&](const T *src, int i0, int i1) {
#pragma unroll
    for (int i = i0; i < i1; i++) {
      const int u = min(grp + 4 * i, NU - 1);
      const T *q = src + uinf[u].x;
      pre_re[i] = q[0];
      pre_im[i] = q[1];
    }
  };
  auto deposit = [&](int i0, int i1) {
#pragma unroll
    for (int i = i0; i < i1; i++) {
      const int u = min(grp + 4 * i, NU - 1);
      const int4 inf = uinf[u];
      const T2 e = cs[inf.y];
      R[inf.z * 16 + col] = e.x * pre_re[i] - e.y * pre_im[i];
      R[inf.w * 16 + col] = e.y * pre_re[i] + e.x * pre_im[i];
    }
  };
#pragma unroll
  for (int i = 0; i < NAU; i++) acc_re[i] = acc_im[i] = T(0);
  if (MODE == 1) load(pvalid ? in + p * (uint64_t)C : zero, 0, NAU);

  for (int c = 0; c < 8; c++) {
    const uint64_t cb = 8 * p + c;
    const bool v = pvalid && (MODE == 0 ? cend[cb] > cstart[cb] : (!cflag || cflag[cb] != 0));
    if (!__any(v)) continue;
    const int th = MODE == 0 ? (c >> 2) & 1 : 1 - ((c >> 2) & 1);
    if (lane <= N) {
      const T *t = CS + ((size_t)c * (N + 1) + lane) * 2;
      cs[lane] = mk2<T>(t[0], t[1]);
    }
    if (MODE == 0) {
      const T *src = v ? in + cb * (uint64_t)C : zero;
#pragma unroll
      for (int i0 = 0; i0 < NAU; i0 += 8) {
        const int i1 = i0 + 8 < NAU ? i0 + 8 : NAU;
        load(src, i0, i1);
        deposit(i0, i1);
      }
    } else {
      deposit(0, NAU);
    }
    V4 y[NB];
    stage<T, NB, KT>(BN, A1 + (size_t)th * KT * 64, R, lane, grp, col, y);
#pragma unroll
    for (int b = 0; b < NB; b++) {
      const int4 d = tmap1[4 * b + grp];
      R[d.x * 16 + col] = y[b][0];
      R[d.y * 16 + col] = y[b][1];
      R[d.z * 16 + col] = y[b][2];
      R[d.w * 16 + col] = y[b][3];
    }
    stage<T, NB, KT>(BN, A2, R, lane, grp, col, y);
#pragma unroll
    for (int b = 0; b < NB; b++) {
      const int4 d = tmap2[4 * b + grp];
      R[d.x * 16 + col] = y[b][0];
      R[d.y * 16 + col] = y[b][1];
      R[d.z * 16 + col] = y[b][2];
      R[d.w * 16 + col] = y[b][3];
    }
    stage<T, NB, KT>(BN, A3 + (size_t)th * KT * 64, R, lane, grp, col, y);
#pragma unroll
    for (int b = 0; b < NB; b++) {
      const int4 d = tmap3[4 * b + grp];
      R[d.x * 16 + col] = y[b][0];
      R[d.y * 16 + col] = y[b][1];
      R[d.z * 16 + col] = y[b][2];
      R[d.w * 16 + col] = y[b][3];
    }
    if (MODE == 0) {
#pragma unroll
      for (int i = 0; i < NAU; i++) {
        const int u = min(grp + 4 * i, NU - 1);
        const T2 e = cs[uinf[u].y];
        const T tr = R[u * 16 + col], tq = R[(NU + u) * 16 + col];
        acc_re[i] += e.x * tr + e.y * tq;
        acc_im[i] += e.x * tq - e.y * tr;
      }
    } else if (v) {
      T *dst = out + cb * (uint64_t)C;
#pragma unroll
      for (int i = 0; i < NAU; i++) {
        const int u = grp + 4 * i;
        if (u < NU) {
          const int4 inf = uinf[u];
          const T2 e = cs[inf.y];
          const T tr = R[u * 16 + col], tq = R[(NU + u) * 16 + col];
          dst[inf.x] += e.x * tr + e.y * tq;
          if (inf.y) dst[inf.x + 1] += e.x * tq - e.y * tr;
        }
      }
    }
  }
  if (MODE == 0 && pvalid) {
    T *dst = out + p * (uint64_t)C;
#pragma unroll
    for (int i = 0; i < NAU; i++) {
      const int u = grp + 4 * i;
      if (u < NU) {
        const int4 inf = uinf[u];
        dst[inf.x] = acc_re[i];
        if (inf.y) dst[inf.x + 1] = acc_im[i];
      }
    }
  }
}

// ---------------------------------------------------------------------
// M2M / L2L for FP64 orders 8..24: the blocks of more than 16 rows need
// more registers than one wavefront holds, and below that several
// wavefronts hide each other's waits, so W wavefronts (4 up to order 15,
// 8 above) share the LDS buffer of 16 parents as in the M2L (wavefront w computes the bins
// b = w (mod W) and owns the units u = grp + 4 w + 4 W i).  The
// directions, tiles and sums are those of k_lap_shift_mfma; the children
// of a parent are added in child order.
// ---------------------------------------------------------------------
template <int N, int W, int WI, int MODE>
__device__ __forceinline__ void lap_shift_loop(
    double *R, const int4 *tmap1, const int4 *tmap2, const int4 *tmap3, const int4 *uinf,
    const double2 (*cs)[N + 1], const unsigned *cmask, unsigned anyc, int lane, int col,
    int grp, bool pvalid, uint64_t p, const double *__restrict__ in, double *__restrict__ out,
    const double *__restrict__ A1, const double *__restrict__ A3,
    const double *__restrict__ A2, const double *__restrict__ zero) {
  using SH = M2LShape<N>;
  constexpr int C = SH::C, NU = SH::NU, KT = SH::KT;
  constexpr int NAW = (NU + 4 * W - 1) / (4 * W);
  constexpr WaveSeq WS = make_seq(SH::BN, W, WI);
  constexpr int NBW = WS.nb;
  const int ubase = grp + 4 * WI;

  double pre_re[NAW], pre_im[NAW];
  auto load_sources = [&](const double *src, const int4 *tu, int ub) {
#pragma unroll
    for (int i = 0; i < NAW; i++) {
      const int u = min(ub + 4 * W * i, NU - 1);
      const double *q = src + tu[u].x;
      pre_re[i] = q[0];
      pre_im[i] = q[1];
    }
  };
  // lanes past the last unit write nothing (one writer per LDS slot)
  auto deposit = [&](const int4 *tu, int ub, int cl, int c) {
#pragma unroll
    for (int i = 0; i < NAW; i++) {
      if (ub + 4 * W * i >= NU) continue;
      const int4 inf = tu[ub + 4 * W * i];
      const double2 e = cs[c][inf.y];
      R[inf.z * 16 + cl] = e.x * pre_re[i] - e.y * pre_im[i];
      R[inf.w * 16 + cl] = e.y * pre_re[i] + e.x * pre_im[i];
    }
  };
  auto write_out = [&](const int4 *tm, const d4 *y, int gp, int cl) {
#pragma unroll
    for (int lb = 0; lb < NBW; lb++) {
      const int4 d = tm[4 * (WI + W * lb) + gp];
      R[d.x * 16 + cl] = y[lb][0];
      R[d.y * 16 + cl] = y[lb][1];
      R[d.z * 16 + cl] = y[lb][2];
      R[d.w * 16 + cl] = y[lb][3];
    }
  };

  double acc_re[NAW], acc_im[NAW];
#pragma unroll
  for (int i = 0; i < NAW; i++) acc_re[i] = acc_im[i] = 0.0;
  if (MODE == 1) load_sources(pvalid ? in + p * (uint64_t)C : zero, uinf, ubase);

  for (int c = 0; c < 8; c++) {
    if (!((anyc >> c) & 1u)) continue;  // the same in every lane of the block
    int z0;
    asm volatile("s_mov_b32 %0, 0" : "=s"(z0));
    const int4 *t1 = tmap1 + z0, *t2 = tmap2 + z0, *t3 = tmap3 + z0, *tu = uinf + z0;
    const int ln = lane + z0, cl = ln & 15, gp = ln >> 4, ub = gp + 4 * WI;
    const bool v = (cmask[cl] >> c) & 1u;
    const int th = MODE == 0 ? (c >> 2) & 1 : 1 - ((c >> 2) & 1);
    if (MODE == 0) load_sources(v ? in + (8 * p + c) * (uint64_t)C : zero, tu, ub);
    __syncthreads();  // every read of R by the last direction done
    deposit(tu, ub, cl, c);
    __syncthreads();

    d4 y[NBW > 0 ? NBW : 1];
    stage_w<double, N, W, WI>(A1 + (size_t)th * KT * 64, R, ln, gp, cl, y);
    __syncthreads();
    write_out(t1, y, gp, cl);
    __syncthreads();
    stage_w<double, N, W, WI>(A2, R, ln, gp, cl, y);
    __syncthreads();
    write_out(t2, y, gp, cl);
    __syncthreads();
    stage_w<double, N, W, WI>(A3 + (size_t)th * KT * 64, R, ln, gp, cl, y);
    __syncthreads();
    write_out(t3, y, gp, cl);
    __syncthreads();

    double *dst = out + (8 * p + c) * (uint64_t)C;
#pragma unroll
    for (int i = 0; i < NAW; i++) {
      const int u = min(ub + 4 * W * i, NU - 1);
      const int4 inf = tu[u];
      const double2 e = cs[c][inf.y];
      const double tr = R[u * 16 + cl], tq = R[(NU + u) * 16 + cl];
      const double orr = e.x * tr + e.y * tq, oi = e.x * tq - e.y * tr;
      if (MODE == 0) {
        acc_re[i] += orr;
        acc_im[i] += oi;
      } else if (v && ub + 4 * W * i < NU) {
        dst[inf.x] += orr;
        if (inf.y) dst[inf.x + 1] += oi;
      }
    }
  }
  if (MODE == 0 && pvalid) {
    double *dst = out + p * (uint64_t)C;
#pragma unroll
    for (int i = 0; i < NAW; i++) {
      const int u = ubase + 4 * W * i;
      if (u >= NU) continue;
      const int4 inf = uinf[u];
      dst[inf.x] = acc_re[i];
      if (inf.y) dst[inf.x + 1] = acc_im[i];
    }
  }
}

template <int N, int W, int MODE>
__global__ __launch_bounds__(64 * W) void k_lap_shift_mfma2(
    const uint32_t *__restrict__ parents, unsigned npar, const uint32_t *__restrict__ cstart,
    const uint32_t *__restrict__ cend, const uint32_t *__restrict__ cflag,
    const double *__restrict__ in, double *__restrict__ out,
    const wbfmm_m2l_plan_t *__restrict__ P, const double *__restrict__ A1,
    const double *__restrict__ A3, const double *__restrict__ A2,
    const double *__restrict__ CS, const double *__restrict__ zero) {
  using SH = M2LShape<N>;
  constexpr int NB = SH::NB, NU = SH::NU, S = SH::S;
  __shared__ double R[(S + 1) * 16];
  __shared__ int4 tmap1[4 * NB], tmap2[4 * NB], tmap3[4 * NB];
  __shared__ int4 uinf[NU];
  __shared__ double2 cs[8][N + 1];
  __shared__ unsigned cmask[16];
  __shared__ unsigned anyc;

  const int tid = threadIdx.x, lane = tid & 63, col = lane & 15, grp = lane >> 4;
  const int wv = __builtin_amdgcn_readfirstlane(tid >> 6);
  for (int i = tid; i < 4 * NB; i += 64 * W) {
    const int b = i >> 2, g = i & 3;
    int t1[4], t2[4], t3[4];
    for (int q = 0; q < 4; q++) {
      const int row = 16 * b + MF<double>::drow(g, q);
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
    uinf[u] = make_int4(n * n + (m ? 2 * m - 1 : 0), m, P->s1[u] & 0xffff,
                        (P->s1[u] >> 16) & 0xffff);
  }
  for (int i = tid; i < (S + 1) * 16; i += 64 * W) R[i] = 0.0;
  for (int i = tid; i < 8 * (N + 1); i += 64 * W)
    cs[i / (N + 1)][i % (N + 1)] = make_double2(CS[2 * i], CS[2 * i + 1]);
  const unsigned pi = blockIdx.x * 16 + col;
  const bool pvalid = pi < npar;
  const uint64_t p = pvalid ? parents[pi] : 0;
  if (tid < 16) {
    const unsigned pj = blockIdx.x * 16 + tid;
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
    for (int j = 0; j < 16; j++) a |= cmask[j];
    anyc = a;
  }
  __syncthreads();
  const unsigned any = __builtin_amdgcn_readfirstlane(anyc);

#define LAP_SHIFT(WI)                                                                       \
  lap_shift_loop<N, W, WI, MODE>(R, tmap1, tmap2, tmap3, uinf, cs, cmask, any, lane, col,  \
                                 grp, pvalid, p, in, out, A1, A3, A2, zero)
  static_assert(W == 4 || W == 8, "four or eight wavefronts per block");
  if constexpr (W == 4) {
    switch (wv) {
      case 0: LAP_SHIFT(0); break;
      case 1: LAP_SHIFT(1); break;
      case 2: LAP_SHIFT(2); break;
      default: LAP_SHIFT(3); break;
    }
  } else {
    switch (wv) {
      case 0: LAP_SHIFT(0); break;
      case 1: LAP_SHIFT(1); break;
      case 2: LAP_SHIFT(2); break;
      case 3: LAP_SHIFT(3); break;
      case 4: LAP_SHIFT(4); break;
      case 5: LAP_SHIFT(5); break;
      case 6: LAP_SHIFT(6); break;
      default: LAP_SHIFT(7); break;
    }
  }
#undef LAP_SHIFT
}

// wavefronts per 16-target chunk of the M2L kernel: 8 for FP64 orders
// 9..15 and 4 otherwise, or 4 / 2 / 1 when WBFMM_M2L_WAVES says so (kept
// for comparison; the results agree to rounding).  The sparse-level
// kernel takes 8 wavefronts per 32 targets unless WBFMM_PACKED_WAVES=4.
static int m2l_waves() {
  static int w = 0;
  if (!w) {
    const char *e = getenv("WBFMM_M2L_WAVES");
    w = e ? atoi(e) : 8;
    if (w != 1 && w != 2 && w != 4) w = 8;
  }
  return w;
}

// ---------------------------------------------------------------------
// Sparse levels (FP64, orders up to 12).  A 16-target chunk runs every
// offset that any of its targets uses; at 40 % box occupancy nearly every
// offset is used by some target of a chunk while each target has a source
// at only 40 % of them, so most columns of each product hold no source.
// Here a block keeps the local expansions of 32 targets in LDS and, for
// each offset, packs the targets whose source box is occupied into
// 16-column tiles (usually one): the products run on occupied pairs only.
// Each target's expansion is the sum of its translations in increasing
// offset order, as in the chunked kernel; the results differ from it by
// rounding only, and do not depend on the launch.
// ---------------------------------------------------------------------
constexpr int kPackT = 32;

template <int N>
struct PackedFits {
  static constexpr bool value =
      N <= 12 && (M2LShape<N>::S + 1) * 16 * 8 + kPackT * M2LShape<N>::C * 8 <= 80 * 1024;
};

template <int N, int W, int WI>
__device__ __forceinline__ void m2l_packed_loop(
    double *R, const int4 *tmap1, const int4 *tmap2, const int4 *tmap3, const int4 *uinf,
    double2 (*cs)[N + 1], const unsigned *anym, const uint8_t (*lists)[kPackT],
    const uint8_t *ncols, const uint32_t *tbox, double *acc, int lane, int px, int py, int pz,
    const wbfmm_m2l_plan_t *__restrict__ P, const double *__restrict__ A1,
    const double *__restrict__ A3, const double *__restrict__ A2,
    const double *__restrict__ CS, const double *__restrict__ zero,
    const double *__restrict__ mps) {
  using SH = M2LShape<N>;
  constexpr int C = SH::C, NU = SH::NU, KT = SH::KT;
  constexpr int NAW = (NU + 4 * W - 1) / (4 * W);
  constexpr WaveSeq WS = make_seq(SH::BN, W, WI);
  constexpr int NBW = WS.nb;
  const int col = lane & 15, ubase = (lane >> 4) + 4 * WI;

  auto offset_of = [&](int k) {
    const int dx = k % 6 - 2 - px, dy = (k / 6) % 6 - 2 - py, dz = k / 36 - 2 - pz;
    return (dx + 3) * 49 + (dy + 3) * 7 + dz + 3;
  };
  // the tiles run in offset order, and within an offset in list order
  auto next_tile = [&](int k, int t0, int &kn, int &tn) {
    if (k >= 0 && k < 216 && t0 + 16 < ncols[k]) {
      kn = k;
      tn = t0 + 16;
      return;
    }
    for (k++; k < 216; k++)
      if ((anym[k >> 5] >> (k & 31)) & 1u) break;
    kn = k;
    tn = 0;
  };
  // every wavefront writes the offset's phases (the same values)
  auto source_of = [&](int k, int t0, int p, int &j) -> const double * {
    if (lane <= N) {
      const double *t = CS + ((size_t)offset_of(k) * (N + 1) + lane) * 2;
      cs[p][lane] = make_double2(t[0], t[1]);
    }
    const int sl = t0 + col;
    j = sl < ncols[k] ? lists[k][sl] : -1;
    if (j < 0) return zero;
    const int dx = k % 6 - 2 - px, dy = (k / 6) % 6 - 2 - py, dz = k / 36 - 2 - pz;
    const uint32_t sb = morton_add(tbox[j], dilate_offset(dx, 0, kMortonX),
                                   dilate_offset(dy, 1, kMortonY), dilate_offset(dz, 2, kMortonZ));
    return mps + (uint64_t)sb * C;
  };
  double pre_re[NAW], pre_im[NAW];
  auto load_sources = [&](const double *src, const int4 *tu, int ub) {
#pragma unroll
    for (int i = 0; i < NAW; i++) {
      const int u = min(ub + 4 * W * i, NU - 1);
      const double *q = src + tu[u].x;
      pre_re[i] = q[0];
      pre_im[i] = q[1];
    }
  };
  // one writer per LDS slot (lanes past the last unit write nothing)
  auto deposit = [&](const int4 *tu, int ub, int cl, int p) {
#pragma unroll
    for (int i = 0; i < NAW; i++) {
      if (ub + 4 * W * i >= NU) continue;
      const int4 inf = tu[ub + 4 * W * i];
      const double2 e = cs[p][inf.y];
      R[inf.z * 16 + cl] = e.x * pre_re[i] - e.y * pre_im[i];
      R[inf.w * 16 + cl] = e.y * pre_re[i] + e.x * pre_im[i];
    }
  };
  auto write_out = [&](const int4 *tm, const d4 *y, int gp, int cl) {
#pragma unroll
    for (int lb = 0; lb < NBW; lb++) {
      const int4 d = tm[4 * (WI + W * lb) + gp];
      R[d.x * 16 + cl] = y[lb][0];
      R[d.y * 16 + cl] = y[lb][1];
      R[d.z * 16 + cl] = y[lb][2];
      R[d.w * 16 + cl] = y[lb][3];
    }
  };

  int k, t0, p = 0, j = -1;
  next_tile(-1, 0, k, t0);
  k = __builtin_amdgcn_readfirstlane(k);
  int th = 0, ri = 0;
  if (k < 216) {
    th = __builtin_amdgcn_readfirstlane(P->th[offset_of(k)]);
    ri = __builtin_amdgcn_readfirstlane(P->rix[offset_of(k)]);
    load_sources(source_of(k, t0, 0, j), uinf, ubase);
    deposit(uinf, ubase, col, 0);
  }
  // stage operands held across stages (StageOps) where they fit
  constexpr bool PF = StageOps<N, W, WI>::fits;
  double av[StageOps<N, W, WI>::n];
  if constexpr (PF) ops_load<double, N, W, WI>(av, A1 + (size_t)th * KT * 64, lane);
  __syncthreads();
  while (k < 216) {
    int z0;
    asm volatile("s_mov_b32 %0, 0" : "=s"(z0));
    const int4 *t1 = tmap1 + z0, *t2 = tmap2 + z0, *t3 = tmap3 + z0, *tu = uinf + z0;
    const int ln = lane + z0, cl = ln & 15, gp = ln >> 4, ub = gp + 4 * WI;
    int kn, tn;
    next_tile(k, t0, kn, tn);
    kn = __builtin_amdgcn_readfirstlane(kn);
    tn = __builtin_amdgcn_readfirstlane(tn);
    int th_n = 0, ri_n = 0, jn = -1;
    if (kn < 216) {
      th_n = __builtin_amdgcn_readfirstlane(P->th[offset_of(kn)]);
      ri_n = __builtin_amdgcn_readfirstlane(P->rix[offset_of(kn)]);
    }

    d4 y[NBW > 0 ? NBW : 1];
    if constexpr (PF)
      stage_wp<double, N, W, WI>(av, A2 + (size_t)ri * KT * 64, R, ln, gp, cl, y);
    else
      stage_w<double, N, W, WI>(A1 + (size_t)th * KT * 64, R, ln, gp, cl, y);
    __syncthreads();
    write_out(t1, y, gp, cl);
    __syncthreads();
    if (kn < 216) load_sources(source_of(kn, tn, p ^ 1, jn), tu, ub);
    if constexpr (PF)
      stage_wp<double, N, W, WI>(av, A3 + (size_t)th * KT * 64, R, ln, gp, cl, y);
    else
      stage_w<double, N, W, WI>(A2 + (size_t)ri * KT * 64, R, ln, gp, cl, y);
    __syncthreads();
    write_out(t2, y, gp, cl);
    __syncthreads();
    if constexpr (PF)
      stage_wp<double, N, W, WI>(av, A1 + (size_t)th_n * KT * 64, R, ln, gp, cl, y);
    else
      stage_w<double, N, W, WI>(A3 + (size_t)th * KT * 64, R, ln, gp, cl, y);
    __syncthreads();
    write_out(t3, y, gp, cl);
    __syncthreads();
    // Zout into the target's expansion in LDS: each (target, unit) has one
    // writer per tile, and a target appears once per offset
    if (j >= 0) {
      double *a = acc + j * C;
#pragma unroll
      for (int i = 0; i < NAW; i++) {
        const int u = ub + 4 * W * i;
        if (u >= NU) continue;
        const int4 inf = tu[u];
        const double2 e = cs[p][inf.y];
        const double tr = R[u * 16 + cl], tq = R[(NU + u) * 16 + cl];
        a[inf.x] += e.x * tr + e.y * tq;
        if (inf.y) a[inf.x + 1] += e.x * tq - e.y * tr;
      }
    }
    __syncthreads();  // every read of R done before the next deposit
    if (kn < 216) deposit(tu, ub, cl, p ^ 1);
    __syncthreads();
    p ^= 1;
    k = kn;
    t0 = tn;
    j = jn;
    th = th_n;
    ri = ri_n;
  }
}

template <int N, int W>
__global__ __launch_bounds__(64 * W) void k_lap_m2l_packed(
    const uint32_t *__restrict__ cls, ClassArgs ca, const uint32_t *__restrict__ occ,
    int level, const double *__restrict__ mps, double *__restrict__ mpr,
    const wbfmm_m2l_plan_t *__restrict__ P, const double *__restrict__ A1,
    const double *__restrict__ A3, const double *__restrict__ A2,
    const double *__restrict__ CS, const double *__restrict__ zero) {
  using SH = M2LShape<N>;
  constexpr int NB = SH::NB, NU = SH::NU, S = SH::S, C = SH::C;
  __shared__ double R[(S + 1) * 16];
  __shared__ double acc[kPackT * C];
  __shared__ int4 tmap1[4 * NB], tmap2[4 * NB], tmap3[4 * NB];
  __shared__ int4 uinf[NU];
  __shared__ double2 cs[2][N + 1];
  __shared__ unsigned vm[kPackT][7];
  __shared__ unsigned anym[7];
  __shared__ uint8_t lists[216][kPackT];
  __shared__ uint8_t ncols[216];
  __shared__ int4 tco[kPackT];
  __shared__ uint32_t tbox[kPackT];

  const int tid = threadIdx.x, lane = tid & 63;
  const int wv = __builtin_amdgcn_readfirstlane(tid >> 6);
  for (int i = tid; i < 4 * NB; i += 64 * W) {
    const int b = i >> 2, g = i & 3;
    int t1[4], t2[4], t3[4];
    for (int q = 0; q < 4; q++) {
      const int row = 16 * b + MF<double>::drow(g, q);
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
    uinf[u] = make_int4(n * n + (m ? 2 * m - 1 : 0), m, P->s1[u] & 0xffff,
                        (P->s1[u] >> 16) & 0xffff);
  }
  for (int i = tid; i < (S + 1) * 16; i += 64 * W) R[i] = 0.0;
  for (int i = tid; i < kPackT * C; i += 64 * W) acc[i] = 0.0;
  for (int i = tid; i < kPackT * 7; i += 64 * W) vm[i / 7][i % 7] = 0u;

  const unsigned nch = ca.chunk_prefix[8], bid = blockIdx.x, q8 = nch / 8;
  const unsigned chunk = bid < 8 * q8 ? (bid & 7) * q8 + (bid >> 3) : bid;
  int c = 0;
  while (c < 7 && chunk >= ca.chunk_prefix[c + 1]) c++;
  const int px = c & 1, py = (c >> 1) & 1, pz = (c >> 2) & 1;
  const int nbox = 1 << level;
  if (tid < kPackT) {
    const unsigned kk = (chunk - ca.chunk_prefix[c]) * kPackT + tid;
    const bool tv = kk < ca.count[c];
    const uint32_t tb = tv ? cls[ca.start[c] + kk] : 0u;
    tbox[tid] = tb;
    tco[tid] = make_int4(compact3(tb), compact3(tb >> 1), compact3(tb >> 2), tv ? 1 : 0);
  }
  __syncthreads();
  for (int idx = tid; idx < kPackT * 216; idx += 64 * W) {
    const int j = idx % kPackT, k = idx / kPackT;
    const int dx = k % 6 - 2 - px, dy = (k / 6) % 6 - 2 - py, dz = k / 36 - 2 - pz;
    if (dx >= -1 && dx <= 1 && dy >= -1 && dy <= 1 && dz >= -1 && dz <= 1) continue;
    const int4 tc = tco[j];
    const int si = tc.x + dx, sj = tc.y + dy, sk = tc.z + dz;
    if (!tc.w || si < 0 || si >= nbox || sj < 0 || sj >= nbox || sk < 0 || sk >= nbox) continue;
    const uint64_t sb = spread3(si) | (spread3(sj) << 1) | (spread3(sk) << 2);
    if ((occ[sb >> 5] >> (sb & 31)) & 1u) atomicOr(&vm[j][k >> 5], 1u << (k & 31));
  }
  __syncthreads();
  if (tid < 7) {
    unsigned a = 0;
    for (int j = 0; j < kPackT; j++) a |= vm[j][tid];
    anym[tid] = a;
  }
  // per offset, the targets with an occupied source, in target order
  for (int k = tid; k < 216; k += 64 * W) {
    int n = 0;
    for (int j = 0; j < kPackT; j++)
      if ((vm[j][k >> 5] >> (k & 31)) & 1u) lists[k][n++] = (uint8_t)j;
    ncols[k] = (uint8_t)n;
  }
  __syncthreads();

#define PACKED(WI)                                                                          \
  m2l_packed_loop<N, W, WI>(R, tmap1, tmap2, tmap3, uinf, cs, anym, lists, ncols, tbox, acc, \
                            lane, px, py, pz, P, A1, A3, A2, CS, zero, mps)
  static_assert(W == 4 || W == 8, "four or eight wavefronts per block");
  if constexpr (W == 4) {
    switch (wv) {
      case 0: PACKED(0); break;
      case 1: PACKED(1); break;
      case 2: PACKED(2); break;
      default: PACKED(3); break;
    }
  } else {
    switch (wv) {
      case 0: PACKED(0); break;
      case 1: PACKED(1); break;
      case 2: PACKED(2); break;
      case 3: PACKED(3); break;
      case 4: PACKED(4); break;
      case 5: PACKED(5); break;
      case 6: PACKED(6); break;
      default: PACKED(7); break;
    }
  }
#undef PACKED
  __syncthreads();
  for (int i = tid; i < kPackT * C; i += 64 * W) {
    const int j = i / C;
    if (tco[j].w) mpr[(size_t)tbox[j] * C + (i - j * C)] += acc[i];
  }
}

// ---------------------------------------------------------------------
// The packed kernel with a double-buffered exchange buffer.  The stages
// alternate between R0 and R1 (R0 -> R1 -> R0 -> R1), so an offset's tile
// needs one barrier per stage and one after the next deposit: 4 instead of
// 8.  The room for the second buffer comes from the local expansions,
// which move from LDS to registers: lane l of wavefront w owns target
// j = l % 32 and the units u = g + 2W i (g = 2w + l / 32), finds its
// target's column in the tile through pos[k][j], and accumulates there.
// The 32 lanes of a half-wavefront read 32 targets' columns of one unit,
// at most 16 distinct columns in distinct banks.  Each target still sums
// its translations in offset order; the results differ from the chunked
// and LDS-accumulator kernels by rounding only.
// ---------------------------------------------------------------------
template <int N>
struct PackedDbFits {
  static constexpr int Sz = (M2LShape<N>::S + 1) * 16 * 8;
  static constexpr bool value = N <= 12 && 2 * Sz + 2 * 216 * kPackT <= 72 * 1024;
};

template <int N, int W, int WI>
__device__ __forceinline__ void m2l_packed_db_loop(
    double *R0, double *R1, const int4 *tmap1, const int4 *tmap2, const int4 *tmap3,
    const int4 *uinf, double2 (*cs)[N + 1], const unsigned *anym,
    const uint8_t (*lists)[kPackT], const uint8_t (*pos)[kPackT], const uint8_t *ncols,
    const int4 *tco, const uint32_t *tbox, const int4 *ktab, const int2 *kth, int lane,
    const double *__restrict__ A1,
    const double *__restrict__ A3, const double *__restrict__ A2,
    const double *__restrict__ CS, const double *__restrict__ zero,
    const double *__restrict__ mps, double *__restrict__ mpr) {
  using SH = M2LShape<N>;
  constexpr int C = SH::C, NU = SH::NU, KT = SH::KT;
  constexpr int NAW = (NU + 4 * W - 1) / (4 * W);
  constexpr int G = 2 * W;                    // unit groups of the accumulators
  constexpr int NAQ = (NU + G - 1) / G;       // accumulator units per lane
  constexpr bool PF = StageOps<N, W, WI>::fits;
  const int col = lane & 15, ubase = (lane >> 4) + 4 * WI;
  const int tj = lane & 31, ug = 2 * WI + (lane >> 5);

  auto next_tile = [&](int k, int t0, int &kn, int &tn) {
    if (k >= 0 && k < 216 && t0 + 16 < ncols[k]) {
      kn = k;
      tn = t0 + 16;
      return;
    }
    for (k++; k < 216; k++)
      if ((anym[k >> 5] >> (k & 31)) & 1u) break;
    kn = k;
    tn = 0;
  };
  auto source_of = [&](int k, int t0, int p) -> const double * {
    const int4 kt = ktab[k];
    if (lane <= N) {
      const double *t = CS + ((size_t)kt.w * (N + 1) + lane) * 2;
      cs[p][lane] = make_double2(t[0], t[1]);
    }
    const int sl = t0 + col;
    const int j = sl < ncols[k] ? lists[k][sl] : -1;
    if (j < 0) return zero;
    const uint32_t sb = morton_add(tbox[j], (uint32_t)kt.x, (uint32_t)kt.y, (uint32_t)kt.z);
    return mps + (uint64_t)sb * C;
  };
  double pre_re[NAW], pre_im[NAW];
  auto load_sources = [&](const double *src, const int4 *tu, int ub) {
#pragma unroll
    for (int i = 0; i < NAW; i++) {
      const int u = min(ub + 4 * W * i, NU - 1);
      const double *q = src + tu[u].x;
      pre_re[i] = q[0];
      pre_im[i] = q[1];
    }
  };
  auto deposit = [&](double *R, const int4 *tu, int ub, int cl, int p) {
#pragma unroll
    for (int i = 0; i < NAW; i++) {
      if (ub + 4 * W * i >= NU) continue;
      const int4 inf = tu[ub + 4 * W * i];
      const double2 e = cs[p][inf.y];
      R[inf.z * 16 + cl] = e.x * pre_re[i] - e.y * pre_im[i];
      R[inf.w * 16 + cl] = e.y * pre_re[i] + e.x * pre_im[i];
    }
  };

  double acc_re[NAQ], acc_im[NAQ];
#pragma unroll
  for (int i = 0; i < NAQ; i++) acc_re[i] = acc_im[i] = 0.0;

  int k, t0, p = 0;
  next_tile(-1, 0, k, t0);
  k = __builtin_amdgcn_readfirstlane(k);
  int th = 0, ri = 0;
  if (k < 216) {
    const int2 tr = kth[k];
    th = __builtin_amdgcn_readfirstlane(tr.x);
    ri = __builtin_amdgcn_readfirstlane(tr.y);
    load_sources(source_of(k, t0, 0), uinf, ubase);
    deposit(R0, uinf, ubase, col, 0);
  }
  double av[StageOps<N, W, WI>::n];
  if constexpr (PF) ops_load<double, N, W, WI>(av, A1 + (size_t)th * KT * 64, lane);
  __syncthreads();
  while (k < 216) {
    int z0;
    asm volatile("s_mov_b32 %0, 0" : "=s"(z0));
    const int4 *t1 = tmap1 + z0, *t2 = tmap2 + z0, *t3 = tmap3 + z0, *tu = uinf + z0;
    const int ln = lane + z0, cl = ln & 15, gp = ln >> 4, ub = gp + 4 * WI;
    int kn, tn;
    next_tile(k, t0, kn, tn);
    kn = __builtin_amdgcn_readfirstlane(kn);
    tn = __builtin_amdgcn_readfirstlane(tn);
    // the next tile's sources are requested before this tile's stages:
    // the registers were emptied by the last deposit, and cs[p ^ 1] was
    // last read before the barrier that ended the previous tile
    int th_n = 0, ri_n = 0;
    if (kn < 216) {
      const int2 tr = kth[kn];
      th_n = __builtin_amdgcn_readfirstlane(tr.x);
      ri_n = __builtin_amdgcn_readfirstlane(tr.y);
      load_sources(source_of(kn, tn, p ^ 1), tu, ub);
    }
    // R0 -> R1 -> R0 -> R1, one barrier per stage
    if constexpr (PF)
      stage_wop<double, N, W, WI>(av, A2 + (size_t)ri * KT * 64, R0, R1, t1, ln, gp, cl);
    else
      stage_wo<double, N, W, WI>(A1 + (size_t)th * KT * 64, R0, R1, t1, ln, gp, cl);
    __syncthreads();
    if constexpr (PF)
      stage_wop<double, N, W, WI>(av, A3 + (size_t)th * KT * 64, R1, R0, t2, ln, gp, cl);
    else
      stage_wo<double, N, W, WI>(A2 + (size_t)ri * KT * 64, R1, R0, t2, ln, gp, cl);
    __syncthreads();
    if constexpr (PF)
      stage_wop<double, N, W, WI>(av, A1 + (size_t)th_n * KT * 64, R0, R1, t3, ln, gp, cl);
    else
      stage_wo<double, N, W, WI>(A3 + (size_t)th * KT * 64, R0, R1, t3, ln, gp, cl);
    __syncthreads();
    // Zout from R1 into the owning lanes' registers; the next deposit
    // goes into R0, which the third stage has finished reading
    {
      const int c = pos[k][tj];
      const int cc = c - t0;
      if (c != 0xff && cc >= 0 && cc < 16) {
#pragma unroll
        for (int i = 0; i < NAQ; i++) {
          const int u = ug + G * i;
          if (u >= NU) continue;
          const int m = tu[u].y;
          const double2 e = cs[p][m];
          const double tr = R1[u * 16 + cc], tq = R1[(NU + u) * 16 + cc];
          acc_re[i] += e.x * tr + e.y * tq;
          acc_im[i] += e.x * tq - e.y * tr;
        }
      }
    }
    if (kn < 216) deposit(R0, tu, ub, cl, p ^ 1);
    __syncthreads();
    p ^= 1;
    k = kn;
    t0 = tn;
    th = th_n;
    ri = ri_n;
  }
  // each lane adds its target's units (m = 0 has no imaginary part)
  if (tco[tj].w) {
    double *out = mpr + (size_t)tbox[tj] * C;
#pragma unroll
    for (int i = 0; i < NAQ; i++) {
      const int u = ug + G * i;
      if (u >= NU) continue;
      const int4 inf = uinf[u];
      out[inf.x] += acc_re[i];
      if (inf.y) out[inf.x + 1] += acc_im[i];
    }
  }
}

// two blocks per CU: at eight wavefronts per block, four per SIMD (at most
// 128 registers each)
template <int N, int W>
__global__ __launch_bounds__(64 * W) __attribute__((amdgpu_waves_per_eu(W == 8 ? 4 : 2)))
void k_lap_m2l_packed_db(
    const uint32_t *__restrict__ cls, ClassArgs ca, const uint32_t *__restrict__ occ,
    int level, const double *__restrict__ mps, double *__restrict__ mpr,
    const wbfmm_m2l_plan_t *__restrict__ P, const double *__restrict__ A1,
    const double *__restrict__ A3, const double *__restrict__ A2,
    const double *__restrict__ CS, const double *__restrict__ zero) {
  using SH = M2LShape<N>;
  constexpr int NB = SH::NB, NU = SH::NU, S = SH::S;
  static_assert(W == 4 || W == 8, "four or eight wavefronts per block");
  __shared__ double R0[(S + 1) * 16], R1[(S + 1) * 16];
  __shared__ int4 tmap1[4 * NB], tmap2[4 * NB], tmap3[4 * NB];
  __shared__ int4 uinf[NU];
  __shared__ double2 cs[2][N + 1];
  __shared__ unsigned vm[kPackT][7];
  __shared__ unsigned anym[7];
  __shared__ uint8_t lists[216][kPackT];
  __shared__ uint8_t pos[216][kPackT];
  __shared__ uint8_t ncols[216];
  __shared__ int4 tco[kPackT];
  __shared__ uint32_t tbox[kPackT];
  __shared__ int4 ktab[216];  // dilated offsets (x, y, z) and offset index
  __shared__ int2 kth[216];   // rotation and coaxial table indices

  const int tid = threadIdx.x, lane = tid & 63;
  const int wv = __builtin_amdgcn_readfirstlane(tid >> 6);
  for (int i = tid; i < 4 * NB; i += 64 * W) {
    const int b = i >> 2, g = i & 3;
    int t1[4], t2[4], t3[4];
    for (int q = 0; q < 4; q++) {
      const int row = 16 * b + MF<double>::drow(g, q);
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
    uinf[u] = make_int4(n * n + (m ? 2 * m - 1 : 0), m, P->s1[u] & 0xffff,
                        (P->s1[u] >> 16) & 0xffff);
  }
  for (int i = tid; i < (S + 1) * 16; i += 64 * W) R0[i] = R1[i] = 0.0;
  for (int i = tid; i < kPackT * 7; i += 64 * W) vm[i / 7][i % 7] = 0u;
  for (int i = tid; i < 216 * kPackT; i += 64 * W) pos[i / kPackT][i % kPackT] = 0xff;

  int c;
  unsigned cj;
  // a block with no chunk leaves before any barrier (uniform over the block)
  if (!block_chunk(blockIdx.x, ca, c, cj)) return;
  const int px = c & 1, py = (c >> 1) & 1, pz = (c >> 2) & 1;
  const int nbox = 1 << level;
  for (int k = tid; k < 216; k += 64 * W) {
    const int dx = k % 6 - 2 - px, dy = (k / 6) % 6 - 2 - py, dz = k / 36 - 2 - pz;
    const int off = (dx + 3) * 49 + (dy + 3) * 7 + dz + 3;
    ktab[k] = make_int4((int)dilate_offset(dx, 0, kMortonX), (int)dilate_offset(dy, 1, kMortonY),
                        (int)dilate_offset(dz, 2, kMortonZ), off);
    kth[k] = make_int2(P->th[off], P->rix[off]);
  }
  if (tid < kPackT) {
    const unsigned kk = cj * kPackT + tid;
    const bool tv = kk < ca.count[c];
    const uint32_t tb = tv ? cls[ca.start[c] + kk] : 0u;
    tbox[tid] = tb;
    tco[tid] = make_int4(compact3(tb), compact3(tb >> 1), compact3(tb >> 2), tv ? 1 : 0);
  }
  __syncthreads();
  for (int idx = tid; idx < kPackT * 216; idx += 64 * W) {
    const int j = idx % kPackT, k = idx / kPackT;
    const int dx = k % 6 - 2 - px, dy = (k / 6) % 6 - 2 - py, dz = k / 36 - 2 - pz;
    if (dx >= -1 && dx <= 1 && dy >= -1 && dy <= 1 && dz >= -1 && dz <= 1) continue;
    const int4 tc = tco[j];
    const int si = tc.x + dx, sj = tc.y + dy, sk = tc.z + dz;
    if (!tc.w || si < 0 || si >= nbox || sj < 0 || sj >= nbox || sk < 0 || sk >= nbox) continue;
    const uint64_t sb = spread3(si) | (spread3(sj) << 1) | (spread3(sk) << 2);
    if ((occ[sb >> 5] >> (sb & 31)) & 1u) atomicOr(&vm[j][k >> 5], 1u << (k & 31));
  }
  __syncthreads();
  if (tid < 7) {
    unsigned a = 0;
    for (int j = 0; j < kPackT; j++) a |= vm[j][tid];
    anym[tid] = a;
  }
  // per offset, the targets with an occupied source (target order) and
  // each target's column position
  for (int k = tid; k < 216; k += 64 * W) {
    int n = 0;
    for (int j = 0; j < kPackT; j++)
      if ((vm[j][k >> 5] >> (k & 31)) & 1u) {
        lists[k][n] = (uint8_t)j;
        pos[k][j] = (uint8_t)n;
        n++;
      }
    ncols[k] = (uint8_t)n;
  }
  __syncthreads();

#define PACKED_DB(WI)                                                                          \
  m2l_packed_db_loop<N, W, WI>(R0, R1, tmap1, tmap2, tmap3, uinf, cs, anym, lists, pos, ncols, \
                               tco, tbox, ktab, kth, lane, A1, A3, A2, CS, zero, mps, mpr)
  if constexpr (W == 4) {
    switch (wv) {
      case 0: PACKED_DB(0); break;
      case 1: PACKED_DB(1); break;
      case 2: PACKED_DB(2); break;
      default: PACKED_DB(3); break;
    }
  } else {
    switch (wv) {
      case 0: PACKED_DB(0); break;
      case 1: PACKED_DB(1); break;
      case 2: PACKED_DB(2); break;
      case 3: PACKED_DB(3); break;
      case 4: PACKED_DB(4); break;
      case 5: PACKED_DB(5); break;
      case 6: PACKED_DB(6); break;
      default: PACKED_DB(7); break;
    }
  }
#undef PACKED_DB
}

// ---------------------------------------------------------------------
// Sparse levels with the offsets packed by class (k_lap_m2l_packed_cls).
// Offsets with the same dz and the same dx^2 + dy^2 share the polar angle
// and the distance, hence all three stage tables; only the azimuthal phases
// differ, and those are applied per column (Zin in the deposit, Zout in the
// accumulation).  For each class (in the order of the key (dz + 3) 32 +
// dx^2 + dy^2, members in cube-index order) a block of 32 targets lists its
// (target, member) pairs with an occupied source in target-major order and
// deals them round-robin over T = max(ceil(n / 16), most pairs of one
// target) tiles: position p goes to tile p mod T, column p div T.  A
// target's pairs are consecutive positions, at most T of them, so a tile
// holds each target at most once, and the lanes that accumulate a target
// find its column in tile t from (r, q, cnt) = (s mod T, s div T, count),
// s its first position.  Fewer, fuller tiles than one offset per tile (on
// the 1e6-point depth-7 tree 4.19 M instead of 4.93 M at level 7), and
// consecutive tiles of a class read the same tables.  The exchange buffer
// is doubled and the expansions stay in registers, as in
// k_lap_m2l_packed_db.  Each target sums its translations in tile order.
// ---------------------------------------------------------------------
// up to order 12 eight wavefronts (two blocks per CU); orders 13..15 sixteen
// (one block per CU, the larger double buffer; registers stay under 128
// with the units and bins spread over twice the wavefronts)
template <int N>
struct PackedClsFits {
  static constexpr int Sz = (M2LShape<N>::S + 1) * 16 * 8;
  static constexpr bool value = (N <= 12 && 2 * Sz + 16 * 1024 <= 72 * 1024) ||
                                (N >= 13 && N <= 15 && 2 * Sz + 32 * 1024 <= 150 * 1024);
  static constexpr int W = N <= 12 ? 8 : 16;
};

template <int N, int W, int WI>
__device__ __forceinline__ void m2l_packed_cls_loop(
    double *R0, double *R1, const int4 *tmap1, const int4 *tmap2, const int4 *tmap3,
    const int4 *uinf, double2 (*cs)[8][N + 1], int ncls, const uint16_t *gbase,
    const uint16_t *gcount, const uint8_t *gT, const int2 *gtab, const uint8_t (*gmem)[8],
    const uint8_t *gnm, const uint8_t *pairs, const uint16_t (*zinfo)[kPackT],
    const int4 *ktab, const int4 *tco, const uint32_t *tbox, int lane,
    const double *__restrict__ A1, const double *__restrict__ A3, const double *__restrict__ A2,
    const double *__restrict__ CS, const double *__restrict__ zero,
    const double *__restrict__ mps, double *__restrict__ mpr) {
  using SH = M2LShape<N>;
  constexpr int C = SH::C, NU = SH::NU, KT = SH::KT;
  constexpr int NAW = (NU + 4 * W - 1) / (4 * W);
  constexpr int G = 2 * W;                    // unit groups of the accumulators
  constexpr int NAQ = (NU + G - 1) / G;       // accumulator units per lane
  constexpr bool PF = StageOps<N, W, WI>::fits;
  const int col = lane & 15, ubase = (lane >> 4) + 4 * WI;
  const int tj = lane & 31, ug = 2 * WI + (lane >> 5);

  // the tile after (g, t): the next tile of class g, else the first tile
  // of the next class with pairs (ncls when there is none)
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
  // this lane's column of tile (g, t): its source expansion and member
  // index (zeros and member 0 for an empty column); with fill, the class's
  // phases into cs[p] (every wavefront writes the same values)
  auto source_of = [&](int g, int t, int p, bool fill, int &icol) -> const double * {
    const int nm = gnm[g];
    if (fill)
      for (int e = lane; e < nm * (N + 1); e += 64) {
        const int i = e / (N + 1), m = e - i * (N + 1);
        const double *tt = CS + ((size_t)ktab[gmem[g][i]].w * (N + 1) + m) * 2;
        cs[p][i][m] = make_double2(tt[0], tt[1]);
      }
    const int T = gT[g], pos = t + col * T;
    icol = 0;
    if (pos >= gcount[g]) return zero;
    const int pr = pairs[gbase[g] + pos];
    icol = pr >> 5;
    const int4 kt = ktab[gmem[g][icol]];
    const uint32_t sb = morton_add(tbox[pr & 31], (uint32_t)kt.x, (uint32_t)kt.y, (uint32_t)kt.z);
    return mps + (uint64_t)sb * C;
  };
  double pre_re[NAW], pre_im[NAW];
  auto load_sources = [&](const double *src, const int4 *tu, int ub) {
#pragma unroll
    for (int i = 0; i < NAW; i++) {
      const int u = min(ub + 4 * W * i, NU - 1);
      const double *q = src + tu[u].x;
      pre_re[i] = q[0];
      pre_im[i] = q[1];
    }
  };
  auto deposit = [&](double *R, const int4 *tu, int ub, int cl, int p, int ic) {
#pragma unroll
    for (int i = 0; i < NAW; i++) {
      if (ub + 4 * W * i >= NU) continue;
      const int4 inf = tu[ub + 4 * W * i];
      const double2 e = cs[p][ic][inf.y];
      R[inf.z * 16 + cl] = e.x * pre_re[i] - e.y * pre_im[i];
      R[inf.w * 16 + cl] = e.y * pre_re[i] + e.x * pre_im[i];
    }
  };

  double acc_re[NAQ], acc_im[NAQ];
#pragma unroll
  for (int i = 0; i < NAQ; i++) acc_re[i] = acc_im[i] = 0.0;

  int g, t, p = 0, icol = 0, icol_n = 0;
  next_tile(-1, 0, g, t);
  g = __builtin_amdgcn_readfirstlane(g);
  int th = 0, ri = 0;
  if (g < ncls) {
    const int2 tr = gtab[g];
    th = __builtin_amdgcn_readfirstlane(tr.x);
    ri = __builtin_amdgcn_readfirstlane(tr.y);
    load_sources(source_of(g, t, 0, true, icol), uinf, ubase);
    deposit(R0, uinf, ubase, col, 0, icol);
  }
  double av[StageOps<N, W, WI>::n];
  if constexpr (PF) ops_load<double, N, W, WI>(av, A1 + (size_t)th * KT * 64, lane);
  __syncthreads();
  while (g < ncls) {
    int z0;
    asm volatile("s_mov_b32 %0, 0" : "=s"(z0));
    const int4 *t1 = tmap1 + z0, *t2 = tmap2 + z0, *t3 = tmap3 + z0, *tu = uinf + z0;
    const int ln = lane + z0, cl = ln & 15, gp = ln >> 4, ub = gp + 4 * WI;
    int gn, tn;
    next_tile(g, t, gn, tn);
    gn = __builtin_amdgcn_readfirstlane(gn);
    tn = __builtin_amdgcn_readfirstlane(tn);
    // the next tile's sources are requested before this tile's stages:
    // the registers were emptied by the last deposit.  The phase buffer
    // changes with the class: a new class's phases go into cs[p ^ 1],
    // last read (by the previous class) before the barrier that ended the
    // previous tile; the tiles of one class share cs[p]
    const bool newc = gn != g;
    const int pn = newc ? p ^ 1 : p;
    int th_n = 0, ri_n = 0;
    if (gn < ncls) {
      const int2 tr = gtab[gn];
      th_n = __builtin_amdgcn_readfirstlane(tr.x);
      ri_n = __builtin_amdgcn_readfirstlane(tr.y);
      load_sources(source_of(gn, tn, pn, newc, icol_n), tu, ub);
    }
    // R0 -> R1 -> R0 -> R1, one barrier per stage
    if constexpr (PF)
      stage_wop<double, N, W, WI>(av, A2 + (size_t)ri * KT * 64, R0, R1, t1, ln, gp, cl);
    else
      stage_wo<double, N, W, WI>(A1 + (size_t)th * KT * 64, R0, R1, t1, ln, gp, cl);
    __syncthreads();
    if constexpr (PF)
      stage_wop<double, N, W, WI>(av, A3 + (size_t)th * KT * 64, R1, R0, t2, ln, gp, cl);
    else
      stage_wo<double, N, W, WI>(A2 + (size_t)ri * KT * 64, R1, R0, t2, ln, gp, cl);
    __syncthreads();
    if constexpr (PF)
      stage_wop<double, N, W, WI>(av, A1 + (size_t)th_n * KT * 64, R0, R1, t3, ln, gp, cl);
    else
      stage_wo<double, N, W, WI>(A3 + (size_t)th * KT * 64, R0, R1, t3, ln, gp, cl);
    __syncthreads();
    // Zout: this lane's target's column of the tile, if it has one
    {
      const int T = gT[g], z = zinfo[g][tj];
      const int r = z & 15, q = (z >> 4) & 15, cnt = z >> 8;
      int d = t - r;
      if (d < 0) d += T;
      if (d < cnt) {
        const int cc = q + (r + d >= T ? 1 : 0);
        const int ic = pairs[gbase[g] + t + cc * T] >> 5;
#pragma unroll
        for (int i = 0; i < NAQ; i++) {
          const int u = ug + G * i;
          if (u >= NU) continue;
          const int m = tu[u].y;
          const double2 e = cs[p][ic][m];
          const double tr = R1[u * 16 + cc], tq = R1[(NU + u) * 16 + cc];
          acc_re[i] += e.x * tr + e.y * tq;
          acc_im[i] += e.x * tq - e.y * tr;
        }
      }
    }
    if (gn < ncls) deposit(R0, tu, ub, cl, pn, icol_n);
    __syncthreads();
    p = pn;
    g = gn;
    t = tn;
    th = th_n;
    ri = ri_n;
    icol = icol_n;
  }
  (void)icol;
  // each lane adds its target's units (m = 0 has no imaginary part)
  if (tco[tj].w) {
    double *out = mpr + (size_t)tbox[tj] * C;
#pragma unroll
    for (int i = 0; i < NAQ; i++) {
      const int u = ug + G * i;
      if (u >= NU) continue;
      const int4 inf = uinf[u];
      out[inf.x] += acc_re[i];
      if (inf.y) out[inf.x + 1] += acc_im[i];
    }
  }
}

template <int N, int W>
__global__ __launch_bounds__(64 * W) __attribute__((amdgpu_waves_per_eu(W >= 8 ? 4 : 2)))
void k_lap_m2l_packed_cls(
    const uint32_t *__restrict__ cls, ClassArgs ca, const uint32_t *__restrict__ occ,
    int level, const double *__restrict__ mps, double *__restrict__ mpr,
    const wbfmm_m2l_plan_t *__restrict__ P, const double *__restrict__ A1,
    const double *__restrict__ A3, const double *__restrict__ A2,
    const double *__restrict__ CS, const double *__restrict__ zero) {
  using SH = M2LShape<N>;
  constexpr int NB = SH::NB, NU = SH::NU, S = SH::S;
  static_assert(W == 4 || W == 8 || W == 16, "four, eight or sixteen wavefronts per block");
  __shared__ double R0[(S + 1) * 16], R1[(S + 1) * 16];
  __shared__ int4 tmap1[4 * NB], tmap2[4 * NB], tmap3[4 * NB];
  __shared__ int4 uinf[NU];
  __shared__ double2 cs[2][8][N + 1];
  __shared__ unsigned vm[kPackT][7];
  __shared__ uint8_t pairs[kPackT * 189];
  __shared__ uint16_t zinfo[kMaxOffCls][kPackT];  // r | q << 4 | cnt << 8
  __shared__ uint8_t ccnt[kMaxOffCls][kPackT];
  __shared__ uint16_t gbase[kMaxOffCls], gcount[kMaxOffCls];
  __shared__ uint8_t gT[kMaxOffCls], gnm[kMaxOffCls], gmem[kMaxOffCls][8];
  __shared__ int2 gtab[kMaxOffCls];
  __shared__ int4 ktab[216];  // dilated offsets (x, y, z) and offset index
  __shared__ int4 tco[kPackT];
  __shared__ uint32_t tbox[kPackT];

  const int tid = threadIdx.x, lane = tid & 63;
  const int wv = __builtin_amdgcn_readfirstlane(tid >> 6);
  for (int i = tid; i < 4 * NB; i += 64 * W) {
    const int b = i >> 2, g = i & 3;
    int t1[4], t2[4], t3[4];
    for (int q = 0; q < 4; q++) {
      const int row = 16 * b + MF<double>::drow(g, q);
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
    uinf[u] = make_int4(n * n + (m ? 2 * m - 1 : 0), m, P->s1[u] & 0xffff,
                        (P->s1[u] >> 16) & 0xffff);
  }
  for (int i = tid; i < (S + 1) * 16; i += 64 * W) R0[i] = R1[i] = 0.0;
  for (int i = tid; i < kPackT * 7; i += 64 * W) vm[i / 7][i % 7] = 0u;

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
    const int nm = kOffCls[c].nm[g];
    gnm[g] = (uint8_t)nm;
    for (int i = 0; i < 8; i++) gmem[g][i] = kOffCls[c].k[g][i];
    // the members share the polar angle and distance: the first one's
    const int k0 = kOffCls[c].k[g][0];
    const int dx = k0 % 6 - 2 - px, dy = (k0 / 6) % 6 - 2 - py, dz = k0 / 36 - 2 - pz;
    const int off = (dx + 3) * 49 + (dy + 3) * 7 + dz + 3;
    gtab[g] = make_int2(P->th[off], P->rix[off]);
  }
  if (tid < kPackT) {
    const unsigned kk = cj * kPackT + tid;
    const bool tv = kk < ca.count[c];
    const uint32_t tb = tv ? cls[ca.start[c] + kk] : 0u;
    tbox[tid] = tb;
    tco[tid] = make_int4(compact3(tb), compact3(tb >> 1), compact3(tb >> 2), tv ? 1 : 0);
  }
  __syncthreads();
  for (int idx = tid; idx < kPackT * 216; idx += 64 * W) {
    const int j = idx % kPackT, k = idx / kPackT;
    const int dx = k % 6 - 2 - px, dy = (k / 6) % 6 - 2 - py, dz = k / 36 - 2 - pz;
    if (dx >= -1 && dx <= 1 && dy >= -1 && dy <= 1 && dz >= -1 && dz <= 1) continue;
    const int4 tc = tco[j];
    const int si = tc.x + dx, sj = tc.y + dy, sk = tc.z + dz;
    if (!tc.w || si < 0 || si >= nbox || sj < 0 || sj >= nbox || sk < 0 || sk >= nbox) continue;
    const uint64_t sb = spread3(si) | (spread3(sj) << 1) | (spread3(sk) << 2);
    if ((occ[sb >> 5] >> (sb & 31)) & 1u) atomicOr(&vm[j][k >> 5], 1u << (k & 31));
  }
  __syncthreads();
  // pairs per (class, target)
  for (int idx = tid; idx < ncls * kPackT; idx += 64 * W) {
    const int g = idx / kPackT, j = idx % kPackT;
    int n = 0;
    for (int i = 0; i < gnm[g]; i++) {
      const int k = gmem[g][i];
      n += (vm[j][k >> 5] >> (k & 31)) & 1u;
    }
    ccnt[g][j] = (uint8_t)n;
  }
  __syncthreads();
  // per class: pair count, tiles, and each target's (r, q, cnt)
  for (int g = tid; g < ncls; g += 64 * W) {
    int n = 0, mx = 0;
    for (int j = 0; j < kPackT; j++) {
      n += ccnt[g][j];
      mx = max(mx, (int)ccnt[g][j]);
    }
    const int T = n ? max((n + 15) / 16, mx) : 0;
    gcount[g] = (uint16_t)n;
    gT[g] = (uint8_t)T;
    int s = 0;
    for (int j = 0; j < kPackT; j++) {
      const int cn = ccnt[g][j];
      // (a target without pairs may start at s = n = 16 T: stored as 0)
      zinfo[g][j] = (uint16_t)(cn ? ((s % T) | ((s / T) << 4) | (cn << 8)) : 0);
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
  // pairs (target | member << 5), target-major within each class
  for (int idx = tid; idx < ncls * kPackT; idx += 64 * W) {
    const int g = idx / kPackT, j = idx % kPackT;
    if (!ccnt[g][j]) continue;
    int s = gbase[g];
    for (int jj = 0; jj < j; jj++) s += ccnt[g][jj];
    for (int i = 0; i < gnm[g]; i++) {
      const int k = gmem[g][i];
      if ((vm[j][k >> 5] >> (k & 31)) & 1u) pairs[s++] = (uint8_t)(j | (i << 5));
    }
  }
  __syncthreads();

#define PACKED_CLS(WI)                                                                        \
  m2l_packed_cls_loop<N, W, WI>(R0, R1, tmap1, tmap2, tmap3, uinf, cs, ncls, gbase, gcount,   \
                                gT, gtab, gmem, gnm, pairs, zinfo, ktab, tco, tbox, lane, A1, \
                                A3, A2, CS, zero, mps, mpr)
  if constexpr (W == 4) {
    switch (wv) {
      case 0: PACKED_CLS(0); break;
      case 1: PACKED_CLS(1); break;
      case 2: PACKED_CLS(2); break;
      default: PACKED_CLS(3); break;
    }
  } else if constexpr (W == 8) {
    switch (wv) {
      case 0: PACKED_CLS(0); break;
      case 1: PACKED_CLS(1); break;
      case 2: PACKED_CLS(2); break;
      case 3: PACKED_CLS(3); break;
      case 4: PACKED_CLS(4); break;
      case 5: PACKED_CLS(5); break;
      case 6: PACKED_CLS(6); break;
      default: PACKED_CLS(7); break;
    }
  } else {
    switch (wv) {
      case 0: PACKED_CLS(0); break;
      case 1: PACKED_CLS(1); break;
      case 2: PACKED_CLS(2); break;
      case 3: PACKED_CLS(3); break;
      case 4: PACKED_CLS(4); break;
      case 5: PACKED_CLS(5); break;
      case 6: PACKED_CLS(6); break;
      case 7: PACKED_CLS(7); break;
      case 8: PACKED_CLS(8); break;
      case 9: PACKED_CLS(9); break;
      case 10: PACKED_CLS(10); break;
      case 11: PACKED_CLS(11); break;
      case 12: PACKED_CLS(12); break;
      case 13: PACKED_CLS(13); break;
      case 14: PACKED_CLS(14); break;
      default: PACKED_CLS(15); break;
    }
  }
#undef PACKED_CLS
}

// the packed kernel for levels where fewer than 60 % of the boxes hold
// targets and there are at least 1024 chunks; WBFMM_M2L_PACKED=0 or 1
// forces the choice at every level and size (read at each call, for
// comparison)
static int m2l_packed_choice() {
  const char *e = getenv("WBFMM_M2L_PACKED");
  return e ? atoi(e) : -1;
}

// part_ext: caller-provided memory for the split partials (the
// allocation-free entry point), or null to allocate them per call
template <typename T, int N>
static int launch_m2l(const uint32_t *cls, const ClassArgs &ca, unsigned nchunks,
                      const uint32_t *occ, int level, const T *mps,
                      T *mpr, const void *plan, const T *A1,
                      const T *A3, const T *A2, const T *CS,
                      const T *zero, const wbfmm_m2l_plan_t *hplan, T *part_ext) {
  if (!plan_matches<N>(hplan)) {
    fprintf(stderr, "wbfmm (HIP): M2L plan for order %d does not match the kernel's bins\n", N);
    return WBFMM_HIP_ERR_ARG;
  }
  // Few chunks leave most of the chip idle while each wavefront works
  // through ~189 offsets in series: split the offsets over up to 8 blocks
  // per chunk.  The split depends on the chunk count only, so a given
  // problem always sums in the same order.
  constexpr int C = (N + 1) * (N + 1);
  unsigned ntot = 0;
  for (int c = 0; c < 8; c++) ntot += ca.count[c];
  // the calling thread's variant (wbfmm_hip_set_variant): the kernel of
  // the sparse levels, or -1 for the choice below
  const int var = wbfmm_hip_get_variant(WBFMM_VARIANT_M2L);
  (void)var;
  if constexpr (sizeof(T) == 8 && (PackedFits<N>::value || PackedClsFits<N>::value)) {
    const int pc = var == 0 ? 0 : m2l_packed_choice();
    const double fill = (double)ntot / (double)((uint64_t)1 << (3 * level));
    // forced on, the kernel runs at any size (it needs no minimum chunk
    // count); chosen automatically, only where it pays
    if (pc == 1 || (pc == -1 && nchunks >= 1024 && fill < 0.6)) {
      ClassArgs cp = ca;
      unsigned nch = 0;
      for (int c = 0; c < 8; c++) {
        cp.chunk_prefix[c] = nch;
        nch += (ca.count[c] + kPackT - 1) / kPackT;
      }
      cp.chunk_prefix[8] = nch;
      const char *pw = getenv("WBFMM_PACKED_WAVES");
      const char *pd = getenv("WBFMM_PACKED_DB");
      const char *pcl = getenv("WBFMM_PACKED_CLS");
      // offsets packed by class (WBFMM_PACKED_CLS=0: one offset per tile)
      if constexpr (PackedClsFits<N>::value) {
        if (var == 2 || (var == -1 && !(pcl && atoi(pcl) == 0) && !(pd && atoi(pd) == 0))) {
          constexpr int WC = PackedClsFits<N>::W;
          if (WC == 16 || !(pw && atoi(pw) == 4)) {
            WBFMM_LAUNCH("lap_m2l", (k_lap_m2l_packed_cls<N, WC>), dim3(nch), dim3(64 * WC), 0,
                         cls, cp, occ, level, (const double *)mps, (double *)mpr,
                         (const wbfmm_m2l_plan_t *)plan, (const double *)A1,
                         (const double *)A3, (const double *)A2, (const double *)CS,
                         (const double *)zero);
          } else if constexpr (WC == 8) {
            WBFMM_LAUNCH("lap_m2l", (k_lap_m2l_packed_cls<N, 4>), dim3(nch), dim3(256), 0, cls,
                         cp, occ, level, (const double *)mps, (double *)mpr,
                         (const wbfmm_m2l_plan_t *)plan, (const double *)A1,
                         (const double *)A3, (const double *)A2, (const double *)CS,
                         (const double *)zero);
          }
          return 0;
        }
      }
      if constexpr (PackedDbFits<N>::value) {
        if (var >= 1 || !(pd && atoi(pd) == 0)) {
          // class order; WBFMM_M2L_XCD=1 interleaves the 8 classes on each
          // XCD (measured 1.4 % slower on the 1e6-point depth-7 solve)
          unsigned grid = nch;
          const char *xe = getenv("WBFMM_M2L_XCD");
          if (xe && atoi(xe) == 1) {
            unsigned mx = 0;
            for (int c = 0; c < 8; c++)
              mx = std::max(mx, cp.chunk_prefix[c + 1] - cp.chunk_prefix[c]);
            cp.qc = (mx + 7) / 8;
            grid = 64 * cp.qc;
          }
          if (!(pw && atoi(pw) == 4))
            WBFMM_LAUNCH("lap_m2l", (k_lap_m2l_packed_db<N, 8>), dim3(grid), dim3(512), 0, cls,
                         cp, occ, level, (const double *)mps, (double *)mpr,
                         (const wbfmm_m2l_plan_t *)plan, (const double *)A1,
                         (const double *)A3, (const double *)A2, (const double *)CS,
                         (const double *)zero);
          else
            WBFMM_LAUNCH("lap_m2l", (k_lap_m2l_packed_db<N, 4>), dim3(grid), dim3(256), 0, cls,
                         cp, occ, level, (const double *)mps, (double *)mpr,
                         (const wbfmm_m2l_plan_t *)plan, (const double *)A1,
                         (const double *)A3, (const double *)A2, (const double *)CS,
                         (const double *)zero);
          return 0;
        }
      }
      if constexpr (PackedFits<N>::value) {
        if (!(pw && atoi(pw) == 4))
          WBFMM_LAUNCH("lap_m2l", (k_lap_m2l_packed<N, 8>), dim3(nch), dim3(512), 0, cls, cp,
                       occ, level, (const double *)mps, (double *)mpr,
                       (const wbfmm_m2l_plan_t *)plan, (const double *)A1, (const double *)A3,
                       (const double *)A2, (const double *)CS, (const double *)zero);
        else
          WBFMM_LAUNCH("lap_m2l", (k_lap_m2l_packed<N, 4>), dim3(nch), dim3(256), 0, cls, cp,
                       occ, level, (const double *)mps, (double *)mpr,
                       (const wbfmm_m2l_plan_t *)plan, (const double *)A1, (const double *)A3,
                       (const double *)A2, (const double *)CS, (const double *)zero);
        return 0;
      }
      // orders 13..15 with the class-packed kernel switched off: chunked
    }
  }
  const int nsplit = m2l_nsplit(nchunks);
  // class order; WBFMM_M2L_XCD=1 interleaves the 8 classes on each XCD at
  // unsplit levels
  ClassArgs cx = ca;
  unsigned grid = nchunks * nsplit;
  {
    const char *oe = getenv("WBFMM_M2L_CLSORDER");
    cx.ord = oe && atoi(oe) == 1 ? 1u : 0u;
  }
  {
    const char *xe = getenv("WBFMM_M2L_XCD");
    if (nsplit == 1 && xe && atoi(xe) == 1) {
      unsigned mx = 0;
      for (int c = 0; c < 8; c++) mx = std::max(mx, ca.chunk_prefix[c + 1] - ca.chunk_prefix[c]);
      cx.qc = (mx + 7) / 8;
      grid = 64 * cx.qc;
    }
  }
  T *part = part_ext;
  if (nsplit > 1 && !part_ext &&
      dmalloc((void **)&part, sizeof(T) * (size_t)nsplit * ntot * C))
    return WBFMM_HIP_ERR_NOMEM;
  // several wavefronts per chunk: eight for FP64 orders 9..15 (at order 12
  // the 11 bins of a stage take 2 per wavefront instead of 3); four fit
  // three per SIMD without spilling at every order; two fit two per SIMD
  // for FP32 and for FP64 up to 12; orders above 15 (blocks spanning
  // several bins) have the four-wavefront kernel only
  if constexpr (sizeof(T) == 8 && N >= 9 && N <= 15) {
    if (m2l_waves() == 8) {
      WBFMM_LAUNCH("lap_m2l", (k_lap_m2l_mfma2<T, N, 8>), dim3(grid), dim3(512), 0,
                   cls, cx, occ, level, mps, mpr, (const wbfmm_m2l_plan_t *)plan, A1, A3, A2, CS,
                   zero, nsplit, part, ntot);
      goto reduce;
    }
  }
  if (N > 15 || m2l_waves() == 4 || m2l_waves() == 8) {
    WBFMM_LAUNCH(sizeof(T) == 4 ? "lap_m2l_f32" : "lap_m2l", (k_lap_m2l_mfma2<T, N, 4>),
                 dim3(grid), dim3(256), 0, cls,
                 cx, occ, level, mps, mpr, (const wbfmm_m2l_plan_t *)plan,
                 A1, A3, A2, CS, zero, nsplit, part, ntot);
  } else if constexpr (N <= 15) {
    if (m2l_waves() >= 2 && (sizeof(T) == 4 || N <= 12))
      WBFMM_LAUNCH(sizeof(T) == 4 ? "lap_m2l_f32" : "lap_m2l", (k_lap_m2l_mfma2<T, N, 2>),
                   dim3(grid), dim3(128), 0, cls,
                   cx, occ, level, mps, mpr, (const wbfmm_m2l_plan_t *)plan,
                   A1, A3, A2, CS, zero, nsplit, part, ntot);
    else
      WBFMM_LAUNCH(sizeof(T) == 4 ? "lap_m2l_f32" : "lap_m2l", (k_lap_m2l_mfma<T, N>),
                   dim3(grid), dim3(64), 0, cls,
                   cx, occ, level, mps, mpr, (const wbfmm_m2l_plan_t *)plan,
                   A1, A3, A2, CS, zero, nsplit, part, ntot);
  }
reduce:
  if (nsplit > 1) {
    const size_t ne = (size_t)ntot * C;
    WBFMM_LAUNCH("m2l_reduce", k_m2l_reduce<T>, dim3((unsigned)((ne + 255) / 256)), dim3(256), 0,
                 cls, ntot, nsplit, C, (const T *)part, mpr);
    if (!part_ext) dfree(part);
  }
  return 0;
}

template <typename T, int N, int MODE>
static int launch_shift(const uint32_t *parents, unsigned npar, const uint32_t *cstart,
                        const uint32_t *cend, const uint32_t *cflag, const T *in,
                        T *out, const void *plan, const T *A1, const T *A3,
                        const T *A2, const T *CS, const T *zero,
                        const wbfmm_m2l_plan_t *hplan) {
  if (!plan_matches<N>(hplan)) {
    fprintf(stderr, "wbfmm (HIP): plan for order %d does not match the kernel's bins\n", N);
    return WBFMM_HIP_ERR_ARG;
  }
  // FP64 orders 8..15: eight wavefronts per 16 parents (WBFMM_SHIFT_WAVES=4
  // or 1 selects the four- or one-wavefront kernel, for comparison; the
  // results agree to rounding)
  if constexpr (sizeof(T) == 8 && N >= 8 && N <= 15) {
    const char *e = getenv("WBFMM_SHIFT_WAVES");
    if (!(e && (atoi(e) == 4 || atoi(e) == 1))) {
      WBFMM_LAUNCH(MODE == 0 ? "lap_m2m" : "lap_l2l", (k_lap_shift_mfma2<N, 8, MODE>),
                   dim3((npar + 15) / 16), dim3(512), 0, parents, npar, cstart, cend, cflag, in,
                   out, (const wbfmm_m2l_plan_t *)plan, A1, A3, A2, CS, zero);
      return 0;
    }
    if (atoi(e) == 4) {
      WBFMM_LAUNCH(MODE == 0 ? "lap_m2m" : "lap_l2l", (k_lap_shift_mfma2<N, 4, MODE>),
                   dim3((npar + 15) / 16), dim3(256), 0, parents, npar, cstart, cend, cflag, in,
                   out, (const wbfmm_m2l_plan_t *)plan, A1, A3, A2, CS, zero);
      return 0;
    }
  }
  if constexpr (N <= 15) {
    WBFMM_LAUNCH(sizeof(T) == 4 ? (MODE == 0 ? "lap_m2m_f32" : "lap_l2l_f32")
                                : (MODE == 0 ? "lap_m2m" : "lap_l2l"),
                 (k_lap_shift_mfma<T, N, MODE>),
                 dim3((npar + 15) / 16), dim3(64), 0, parents, npar, cstart, cend, cflag, in,
                 out, (const wbfmm_m2l_plan_t *)plan, A1, A3, A2, CS, zero);
  } else if constexpr (sizeof(T) == 8) {
    WBFMM_LAUNCH(MODE == 0 ? "lap_m2m" : "lap_l2l", (k_lap_shift_mfma2<N, 8, MODE>),
                 dim3((npar + 15) / 16), dim3(512), 0, parents, npar, cstart, cend, cflag, in,
                 out, (const wbfmm_m2l_plan_t *)plan, A1, A3, A2, CS, zero);
  } else {
    return WBFMM_HIP_ERR_ORDER;  // FP32 passes: orders up to 15
  }
  return 0;
}


// launch_m2l / launch_shift for the orders n..HI, chosen at run time
template <typename T, int n, int HI>
int m2l_orders(int N, const uint32_t *cls, const ClassArgs &ca, unsigned nch,
               const uint32_t *occ, int level, const T *mps, T *mpr, const void *plan,
               const T *A1, const T *A3, const T *A2, const T *CS, const T *zero,
               const wbfmm_m2l_plan_t *hplan, T *part) {
  if constexpr (n > HI) {
    return WBFMM_HIP_ERR_ORDER;
  } else {
    if (N == n)
      return launch_m2l<T, n>(cls, ca, nch, occ, level, mps, mpr, plan, A1, A3, A2, CS, zero,
                              hplan, part);
    return m2l_orders<T, n + 1, HI>(N, cls, ca, nch, occ, level, mps, mpr, plan, A1, A3, A2,
                                    CS, zero, hplan, part);
  }
}

template <typename T, int n, int HI>
int shift_orders(int mode, int N, const uint32_t *parents, unsigned npar,
                 const uint32_t *cstart, const uint32_t *cend, const uint32_t *cflag,
                 const T *in, T *out, const void *plan, const T *A1, const T *A3, const T *A2,
                 const T *CS, const T *zero, const wbfmm_m2l_plan_t *hplan) {
  if constexpr (n > HI) {
    return WBFMM_HIP_ERR_ORDER;
  } else {
    if (N == n)
      return mode == 0 ? launch_shift<T, n, 0>(parents, npar, cstart, cend, cflag, in, out,
                                               plan, A1, A3, A2, CS, zero, hplan)
                       : launch_shift<T, n, 1>(parents, npar, cstart, cend, cflag, in, out,
                                               plan, A1, A3, A2, CS, zero, hplan);
    return shift_orders<T, n + 1, HI>(mode, N, parents, npar, cstart, cend, cflag, in, out,
                                      plan, A1, A3, A2, CS, zero, hplan);
  }
}

// the four range functions declared in m2l_lap_ranges.h
#define WBFMM_LAP_RANGE_DEFINE(LO, HI)                                                       \
  int lap_m2l_range_##LO##_##HI(int N, const uint32_t *cls, const ClassArgs &ca,            \
                                unsigned nch, const uint32_t *occ, int level,               \
                                const double *mps, double *mpr, const void *plan,           \
                                const double *A1, const double *A3, const double *A2,       \
                                const double *CS, const double *zero,                       \
                                const wbfmm_m2l_plan_t *hplan, double *part) {              \
    return m2l_orders<double, LO, HI>(N, cls, ca, nch, occ, level, mps, mpr, plan, A1, A3,  \
                                      A2, CS, zero, hplan, part);                           \
  }                                                                                         \
  int lap_m2l_range_##LO##_##HI(int N, const uint32_t *cls, const ClassArgs &ca,            \
                                unsigned nch, const uint32_t *occ, int level,               \
                                const float *mps, float *mpr, const void *plan,             \
                                const float *A1, const float *A3, const float *A2,          \
                                const float *CS, const float *zero,                         \
                                const wbfmm_m2l_plan_t *hplan, float *part) {               \
    return m2l_orders<float, LO, HI>(N, cls, ca, nch, occ, level, mps, mpr, plan, A1, A3,   \
                                     A2, CS, zero, hplan, part);                            \
  }                                                                                         \
  int lap_shift_range_##LO##_##HI(int mode, int N, const uint32_t *parents, unsigned npar,  \
                                  const uint32_t *cstart, const uint32_t *cend,             \
                                  const uint32_t *cflag, const double *in, double *out,     \
                                  const void *plan, const double *A1, const double *A3,     \
                                  const double *A2, const double *CS, const double *zero,   \
                                  const wbfmm_m2l_plan_t *hplan) {                          \
    return shift_orders<double, LO, HI>(mode, N, parents, npar, cstart, cend, cflag, in,    \
                                        out, plan, A1, A3, A2, CS, zero, hplan);            \
  }                                                                                         \
  int lap_shift_range_##LO##_##HI(int mode, int N, const uint32_t *parents, unsigned npar,  \
                                  const uint32_t *cstart, const uint32_t *cend,             \
                                  const uint32_t *cflag, const float *in, float *out,       \
                                  const void *plan, const float *A1, const float *A3,       \
                                  const float *A2, const float *CS, const float *zero,      \
                                  const wbfmm_m2l_plan_t *hplan) {                          \
    return shift_orders<float, LO, HI>(mode, N, parents, npar, cstart, cend, cflag, in,     \
                                       out, plan, A1, A3, A2, CS, zero, hplan);             \
  }

}  // namespace wbfmm

#endif
