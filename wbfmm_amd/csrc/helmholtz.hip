// helmholtz.hip -- gfx950 kernels of the Helmholtz FMM path (FP64).
//
// Kernel G = h0(kr)/(4 pi) = e^{ikr}/(4 pi i k r) (src/direct.c:35-134).
// Coefficients are complex, C_n^m for -n <= m <= n at index n(n+1)+m
// (src/wbfmm.h:242), two doubles (re, im) each; level arrays are box-major
// on the device (box b at b * 2 (N+1)^2 doubles).
//
// Translations follow the reference's rotate -> coaxial shift -> rotate
// back sequence (src/pass.c:56-105, src/shift.c:244-585) with the
// Gumerov-Duraiswami H(theta) coefficients (src/rotations.c:131-207) and
// coaxial tables (E|F)_{l n}^m (src/translations.c:205-604):
//   rotate:  Co_n^nu = e^{-i nu phi} sum_m H_n^{nu,m}(theta) e^{i m chi} Ci_n^m
//   coaxial: Co_n^m  = sum_l T(n, l, |m|) Ci_l^m,
//            T = E_{l n}^{|m|} (l >= n), (-1)^{n+l} E_{n l}^{|m|} (l < n).
// The inverse rotation swaps phi and chi.
//
// Kernel shapes (first version, one wavefront per box, vectors in LDS):
//   P2M      one wave per leaf box; per point the regular functions
//            R_n^m = j_n(kr) Pbar_n^|m| e^{i m phi} go to LDS (lane m
//            computes column m) and lane m accumulates orders +-m
//   M2M/L2L  one wave per parent, the 8 children in turn
//   M2L      one wave per target box, its list-4 sources in turn, the
//            local expansion accumulated in LDS
//   L2P+P2P  one thread per target, regular functions streamed by order
//            with the gradient in the re-indexed form of G&D (2004) eq 3.7

#include <cstdio>
#include <cstring>

#include <hip/hip_runtime.h>

#include <mutex>

#include "hip_common.h"
#include "m2l_mfma.h"

namespace wbfmm {

__device__ __forceinline__ int hidx(int n, int m) { return n * (n + 1) + m; }
__device__ __forceinline__ int ridx(int nu, int m, int n) {
  return n * (4 * n - 1) * (n + 1) / 6 + (nu + n) * (n + 1) + m;
}
__device__ __forceinline__ int cxi(int l, int m, int n) {
  return l * (l + 1) * (l + 2) / 6 + n * (n + 1) / 2 + m;
}

__device__ __forceinline__ double2 cmul(double2 a, double2 b) {
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ double2 cmulc(double2 a, double2 b) {  // a * conj(b)
  return make_double2(a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y);
}
__device__ __forceinline__ double2 cadd(double2 a, double2 b) {
  return make_double2(a.x + b.x, a.y + b.y);
}
__device__ __forceinline__ double2 cscale(double2 a, double s) {
  return make_double2(a.x * s, a.y * s);
}

// Gumerov-Duraiswami recursion coefficients a_n^m, b_n^m (G&D 2003)
__device__ __forceinline__ double rec_a(int n, int m) {
  m = m < 0 ? -m : m;
  if (n < m || n < 0) return 0.0;
  return sqrt((double)(n + 1 + m) * (n + 1 - m) / ((double)(2 * n + 1) * (2 * n + 3)));
}
__device__ __forceinline__ double rec_b(int n, int m) {
  if (n < 0 || n < (m < 0 ? -m : m)) return 0.0;
  const double v = sqrt((double)(n - m - 1) * (n - m) / ((double)(2 * n - 1) * (2 * n + 1)));
  return m >= 0 ? v : -v;
}

// j_n(x), n = 0..N into j[]: upward recursion while x > n+1, the power
// series otherwise (src/util.c:118-175)
__device__ inline double bessel_j_series(int n, double x) {
  double sc = 1.0 / (2 * n + 1), term = 1.0, s = 1.0;
  const double x2 = x * x;
  for (int q = 0; q < n; q++) sc *= x / (2 * q + 1);
  for (int q = 0; q < 80 && fabs(term) > 1e-17; q++) {
    term *= -x2 / ((2 * q + 2) * (double)(2 * n + 2 * (q + 1) + 1));
    s += term;
  }
  return s * sc;
}

__device__ inline void bessel_j(int N, double x, double *j) {
  if (x > 0.0) {
    j[0] = sin(x) / x;
    if (N >= 1) j[1] = j[0] / x - cos(x) / x;
  } else {
    j[0] = 1.0;
    if (N >= 1) j[1] = 0.0;
  }
  for (int n = 1; n < N; n++)
    j[n + 1] = (x > n + 1) ? j[n] * (2 * n + 1) / x - j[n - 1] : bessel_j_series(n + 1, x);
}

// j_n(x) of one order by the rule of bessel_j (upward recursion from j_0,
// j_1 for x > n, the power series otherwise), so that the lanes of a
// wavefront compute the orders side by side
__device__ inline double bessel_j_one(int n, double x) {
  if (!(x > 0.0)) return n == 0 ? 1.0 : 0.0;
  const double j0 = sin(x) / x;
  if (n == 0) return j0;
  const double j1 = j0 / x - cos(x) / x;
  if (n == 1) return j1;
  if (!(x > n)) return bessel_j_series(n, x);
  double a = j0, b = j1;
  for (int q = 1; q < n; q++) {
    const double c = b * (2 * q + 1) / x - a;
    a = b;
    b = c;
  }
  return b;
}

// spherical coordinates of x - x0: r, cos/sin theta, cos/sin phi
__device__ __forceinline__ void sph(double dx, double dy, double dz, double &r,
                                    double &C, double &S, double &cp, double &sp) {
  r = sqrt(dx * dx + dy * dy + dz * dz);
  if (r == 0.0) {
    C = 1.0; S = 0.0; cp = 1.0; sp = 0.0;
    return;
  }
  C = dz / r;
  const double rho = sqrt(dx * dx + dy * dy);
  S = rho / r;
  if (rho == 0.0) {
    cp = 1.0; sp = 0.0;
  } else {
    cp = dx / rho; sp = dy / rho;
  }
}

// Pbar_m^m (G&D normalisation, no Condon-Shortley phase)
__device__ __forceinline__ double pbar_mm(int m, double S) {
  double p = 1.0 / sqrt(4.0 * kPi);
  for (int i = 1; i <= m; i++) p *= sqrt((2.0 * i + 1) / (2.0 * i)) * S;
  return p;
}
// Pbar_{n+1}^m from Pbar_n^m, Pbar_{n-1}^m (src/util.c:81-115)
__device__ __forceinline__ double pbar_up(int n, int m, double C, double pn, double pnm1) {
  const double v = C * sqrt(2.0 * n + 1) * pn - sqrt((double)(n * n - m * m) / (2 * n - 1)) * pnm1;
  return v / sqrt((double)((n + 1) * (n + 1) - m * m) / (2 * n + 3));
}

// The factors of these recurrences per (n, m), as one table per process
// built once on the device by the functions above (the same values), for
// the field kernels: their loops over (n, m) are the same in every lane,
// so the table is read with scalar loads instead of a square root and a
// division per factor and target.
constexpr int kRT = WBFMM_HIP_MAX_ORDER + 3;
struct RecTab {
  double ra[kRT][kRT];          // rec_a(i - 1, m)
  double rb[kRT][2 * kRT + 1];  // rec_b(i, m - kRT)
  double pa[kRT];               // sqrt(2 i + 1)
  double pb[kRT][kRT];          // sqrt((i^2 - m^2) / (2 i - 1))
  double pd[kRT][kRT];          // sqrt(((i + 1)^2 - m^2) / (2 i + 3))
};

__global__ void k_rec_tab(RecTab *T) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < kRT * kRT) {
    const int i = t / kRT, m = t % kRT;
    T->ra[i][m] = rec_a(i - 1, m);
    T->pb[i][m] = sqrt((double)(i * i - m * m) / (2 * i - 1));
    T->pd[i][m] = sqrt((double)((i + 1) * (i + 1) - m * m) / (2 * i + 3));
  }
  if (t < kRT * (2 * kRT + 1)) {
    const int i = t / (2 * kRT + 1), m = t % (2 * kRT + 1) - kRT;
    T->rb[i][m + kRT] = rec_b(i, m);
  }
  if (t < kRT) T->pa[t] = sqrt(2.0 * t + 1);
}

// rec_a(n, m) for n >= -1, |m| <= n + 1 <= WBFMM_HIP_MAX_ORDER + 1
__device__ __forceinline__ double tab_a(const RecTab *__restrict__ T, int n, int m) {
  return T->ra[n + 1][m < 0 ? -m : m];
}
// rec_b(n, m) for 0 <= n <= WBFMM_HIP_MAX_ORDER + 2, |m| <= n + 1
__device__ __forceinline__ double tab_b(const RecTab *__restrict__ T, int n, int m) {
  return T->rb[n][m + kRT];
}
// pbar_up with the table's factors (0 <= m <= n <= WBFMM_HIP_MAX_ORDER + 1)
__device__ __forceinline__ double tab_pbar_up(const RecTab *__restrict__ T, int n, int m, double C,
                                              double pn, double pnm1) {
  const double v = C * T->pa[n] * pn - T->pb[n][m] * pnm1;
  return v / T->pd[n][m];
}

// ---------------------------------------------------------------------
// wave-cooperative operators on LDS vectors (double2, index hidx)
// ---------------------------------------------------------------------

// out (+)= e^{-i nu ph} sum_m H e^{i m ch} in, degrees 0..N; tmp: (N+1)^2,
// ep/ec: N+1 phase slots
__device__ void rotate_wave(const double2 *in, double2 *out, double2 *tmp,
                            double2 *ep, double2 *ec, int N,
                            const double *__restrict__ H, double ph,
                            double ch, bool add) {
  const int lane = threadIdx.x;
  if (lane <= N) {
    double s, c;
    sincos((double)lane * ch, &s, &c);
    ec[lane] = make_double2(c, s);
    sincos(-(double)lane * ph, &s, &c);
    ep[lane] = make_double2(c, s);
  }
  __syncthreads();
  const int nc = (N + 1) * (N + 1);
  for (int r = lane; r < nc; r += kWave) {
    const int n = isqrt_deg(r), m = r - n * (n + 1);
    const double2 e = m >= 0 ? ec[m] : make_double2(ec[-m].x, -ec[-m].y);
    tmp[r] = cmul(e, in[r]);
  }
  __syncthreads();
  for (int r = lane; r < nc; r += kWave) {
    const int n = isqrt_deg(r), nu = r - n * (n + 1);
    double2 acc = make_double2(0.0, 0.0);
    for (int m = -n; m <= n; m++) {
      const double h = m >= 0 ? H[ridx(nu, m, n)] : H[ridx(-nu, -m, n)];
      const double2 v = tmp[hidx(n, m)];
      acc.x += h * v.x;
      acc.y += h * v.y;
    }
    const double2 e = nu >= 0 ? ep[nu] : make_double2(ep[-nu].x, -ep[-nu].y);
    acc = cmul(e, acc);
    out[r] = add ? cadd(out[r], acc) : acc;
  }
  __syncthreads();
}

// out (+)= coaxial translation of in (degrees 0..Ni) to degrees 0..No
__device__ void coax_wave(const double2 *in, int Ni, double2 *out, int No,
                          const double2 *__restrict__ E, bool add) {
  const int lane = threadIdx.x, nc = (No + 1) * (No + 1);
  for (int r = lane; r < nc; r += kWave) {
    const int n = isqrt_deg(r), m = r - n * (n + 1), am = m < 0 ? -m : m;
    double2 acc = make_double2(0.0, 0.0);
    for (int l = am; l <= Ni; l++) {
      const int idx = l >= n ? cxi(l, am, n) : cxi(n, am, l);
      double2 e = E[idx];
      if (l < n && ((n + l) & 1)) e = make_double2(-e.x, -e.y);
      acc = cadd(acc, cmul(e, in[hidx(l, m)]));
    }
    out[r] = add ? cadd(out[r], acc) : acc;
  }
  __syncthreads();
}

// ---------------------------------------------------------------------
// P2M (src/boxes.c:341-490, src/expansions.c:52-123, :545-631)
// ---------------------------------------------------------------------
// monopole:  M_n^m += q conj(R_n^m(y - c))
// dipole:    M_n^m += -d conj(nrm . grad R_n^m(y - c))  (the source
//            d nrm.grad_x G, src/local.c:155-166)
// A block of 256 threads per box takes the box's points in batches of pb
// (<= 8, set by the LDS the batch needs):
//   (1) j_n(k r_p) for every point p of the batch and order n, one thread
//       each;
//   (2) column m of R_n^m(p) (n = m..N+1, m >= 0) for every (p, m), one
//       thread each, into LDS;
//   (3) every thread adds the batch's contributions, point by point in
//       source order, to the coefficients n(n+1)+m it owns (tid + 256 i),
//       held in registers until the box is done.
// The sums run in source order per coefficient: results do not depend on
// the launch.
constexpr int kP2MThreads = 256, kP2MBatch = 8;
constexpr int kP2MCoef =
    ((WBFMM_HIP_MAX_ORDER + 1) * (WBFMM_HIP_MAX_ORDER + 1) + kP2MThreads - 1) / kP2MThreads;

// O: double2 (FP64 coefficients) or float2 (FP32 coefficients, each degree
// n divided by sc[n] in FP64 before rounding; sc NULL: no scaling)
// NC: coefficients per thread, ceil((N+1)^2 / 256); three blocks per CU
// (the LDS of a batch is at most 60 KiB)
template <typename O, int NC>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3))) void k_hel_p2m(
    const uint32_t *__restrict__ boxes, unsigned nbox,
    const uint32_t *__restrict__ start, const uint32_t *__restrict__ end,
    const double *__restrict__ xyz, const double2 *__restrict__ q,
    const double *__restrict__ nrm, const double2 *__restrict__ dip,
    unsigned level, double x0, double y0, double z0, double D, int N,
    double k, int pb, const double *__restrict__ sc, O *__restrict__ mps,
    const RecTab *__restrict__ RT) {
  extern __shared__ double2 smem[];
  const int N1 = N + 1, tid = threadIdx.x, NT = (N1 + 1) * (N1 + 2) / 2;
  double2 *R = smem;                        // pb x NT: R_n^m, n <= N+1, m >= 0
  double *jn = (double *)(R + (size_t)pb * NT);  // pb x (N+2)
  const unsigned bi = blockIdx.x;
  if (bi >= nbox) return;
  const uint32_t b = boxes[bi];
  const double w = D / (double)(1u << level);
  const double xc = x0 + (compact3(b) + 0.5) * w, yc = y0 + (compact3(b >> 1) + 0.5) * w,
               zc = z0 + (compact3(b >> 2) + 0.5) * w;
  const uint32_t s0 = start[b], s1 = end[b];
  double2 acc[NC];
#pragma unroll
  for (int i = 0; i < NC; i++) acc[i] = make_double2(0.0, 0.0);

  for (uint32_t sb = s0; sb < s1; sb += (uint32_t)pb) {
    const int np = (int)min((uint32_t)pb, s1 - sb);
    for (int it = tid; it < np * (N1 + 1); it += kP2MThreads) {
      const int p = it / (N1 + 1), n = it - p * (N1 + 1);
      const uint32_t s = sb + p;
      double r, C, S, cp, sp;
      sph(xyz[4 * s] - xc, xyz[4 * s + 1] - yc, xyz[4 * s + 2] - zc, r, C, S, cp, sp);
      jn[p * (N1 + 1) + n] = bessel_j_one(n, k * r);
    }
    __syncthreads();
    for (int it = tid; it < np * (N1 + 1); it += kP2MThreads) {
      const int p = it / (N1 + 1), m = it - p * (N1 + 1);
      const uint32_t s = sb + p;
      double r, C, S, cp, sp;
      sph(xyz[4 * s] - xc, xyz[4 * s + 1] - yc, xyz[4 * s + 2] - zc, r, C, S, cp, sp);
      double sn, cs;
      sincos((double)m * atan2(sp, cp), &sn, &cs);
      const double *j = jn + p * (N1 + 1);
      double2 *Rp = R + (size_t)p * NT;
      double pv = pbar_mm(m, S), pprev = 0.0;
      for (int n = m; n <= N1; n++) {
        Rp[n * (n + 1) / 2 + m] = make_double2(j[n] * pv * cs, j[n] * pv * sn);
        const double pn = (n + 1 <= N1) ? tab_pbar_up(RT, n, m, C, pv, pprev) : 0.0;
        pprev = pv;
        pv = pn;
      }
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < NC; i++) {
      const int c = tid + kP2MThreads * i;
      if (c >= N1 * N1) break;
      const int n = isqrt_deg(c), mm = c - n * (n + 1);
      // G&D 2004 eq 3.7 coefficients of this (n, mm)
      const double a0 = tab_a(RT, n - 1, mm), a1 = -tab_a(RT, n, mm);
      const double bp0 = tab_b(RT, n + 1, -mm - 1), bp1 = -tab_b(RT, n, mm);
      const double bm0 = tab_b(RT, n + 1, mm - 1), bm1 = -tab_b(RT, n, -mm);
      // the LDS slots of the R_{n'}^{m'} this coefficient reads, with
      // R^{-m'} = conj R^{m'} as a sign on the imaginary part and a weight 0
      // (slot 0) for (n', m') outside the table: (n, mm), (n - 1, mm),
      // (n + 1, mm), (n + 1, mm + 1), (n - 1, mm + 1), (n + 1, mm - 1),
      // (n - 1, mm - 1)
      int ix[7];
      double sx[7], sy[7];
      {
        const int nqs[7] = {n, n - 1, n + 1, n + 1, n - 1, n + 1, n - 1};
        const int mqs[7] = {mm, mm, mm, mm + 1, mm + 1, mm - 1, mm - 1};
#pragma unroll
        for (int e = 0; e < 7; e++) {
          const int nq = nqs[e], mq = mqs[e];
          const bool ok = !(nq < 0 || nq > N1 || mq > nq || -mq > nq);
          ix[e] = ok ? nq * (nq + 1) / 2 + (mq < 0 ? -mq : mq) : 0;
          sx[e] = ok ? 1.0 : 0.0;
          sy[e] = ok ? (mq >= 0 ? 1.0 : -1.0) : 0.0;
        }
      }
      double2 ac = acc[i];
      for (int p = 0; p < np; p++) {
        const uint32_t s = sb + p;
        const double2 *Rp = R + (size_t)p * NT;
        auto Rg = [&](int e) -> double2 {
          const double2 v = Rp[ix[e]];
          return make_double2(v.x * sx[e], v.y * sy[e]);
        };
        double2 v = make_double2(0.0, 0.0);
        if (q) v = cmulc(q[s], Rg(0));
        if (dip) {
          const double2 dd = dip[s];
          const double nx = nrm[4 * s], ny = nrm[4 * s + 1], nz = nrm[4 * s + 2];
          const double2 dz = cscale(cadd(cscale(Rg(1), a0), cscale(Rg(2), a1)), k);
          const double2 dp = cscale(cadd(cscale(Rg(3), bp0), cscale(Rg(4), bp1)), k);
          const double2 dm = cscale(cadd(cscale(Rg(5), bm0), cscale(Rg(6), bm1)), k);
          // gx = (dp + dm)/2, gy = -i (dp - dm)/2
          const double2 gx = cscale(cadd(dp, dm), 0.5);
          const double2 dd2 = cscale(cadd(dp, cscale(dm, -1.0)), 0.5);
          const double2 gy = make_double2(dd2.y, -dd2.x);
          const double2 g = cadd(cadd(cscale(gx, nx), cscale(gy, ny)), cscale(dz, nz));
          const double2 t = cmulc(dd, g);
          v = cadd(v, make_double2(-t.x, -t.y));
        }
        ac = cadd(ac, v);
      }
      acc[i] = ac;
    }
    __syncthreads();  // R and jn are rewritten by the next batch
  }
  O *out = mps + (size_t)b * N1 * N1;
#pragma unroll
  for (int i = 0; i < NC; i++) {
    const int c = tid + kP2MThreads * i;
    if (c < N1 * N1) {
      const double s = sc ? 1.0 / sc[isqrt_deg(c)] : 1.0;
      O o = out[c];
      o.x += acc[i].x * s;
      o.y += acc[i].y * s;
      out[c] = o;
    }
  }
}

// ---------------------------------------------------------------------
// M2M (src/pass.c:581-632, src/shift.c:244-401): the 8 children of each
// parent, rotated along the child -> parent direction, R|R shift by
// sqrt(3)/2 w_child, rotated back and summed into the parent
// ---------------------------------------------------------------------
__device__ __forceinline__ void child_dir(int c, int sgn, double &th, double &ch) {
  const double ex = sgn * (0.5 - (c & 1)), ey = sgn * (0.5 - ((c >> 1) & 1)),
               ez = sgn * (0.5 - ((c >> 2) & 1));
  th = acos(ez / sqrt(ex * ex + ey * ey + ez * ez));
  ch = atan2(ey, ex);
}

// Hud: H tables of the two polar angles acos(+-1/sqrt 3) (index 0: +, 1: -)
__global__ __launch_bounds__(64) void k_hel_m2m(
    const uint32_t *__restrict__ parents, unsigned npar,
    const uint32_t *__restrict__ cstart, const uint32_t *__restrict__ cend,
    int Nc, int Np, const double2 *__restrict__ mps_c, double2 *__restrict__ mps_p,
    const double *__restrict__ Hud, size_t nrot, const double2 *__restrict__ SS) {
  extern __shared__ double2 smem[];
  const int L = Nc > Np ? Nc : Np, L1 = (L + 1) * (L + 1);
  double2 *in = smem, *t1 = in + L1, *t2 = t1 + L1, *tmp = t2 + L1, *acc = tmp + L1;
  double2 *ep = acc + L1, *ec = ep + L + 1;
  const int lane = threadIdx.x;
  if (blockIdx.x >= npar) return;
  const uint32_t p = parents[blockIdx.x];
  for (int r = lane; r < (Np + 1) * (Np + 1); r += kWave) acc[r] = make_double2(0.0, 0.0);
  for (int c = 0; c < 8; c++) {
    const uint32_t cb = 8 * p + c;
    if (cend[cb] <= cstart[cb]) continue;
    const double2 *src = mps_c + (size_t)cb * (Nc + 1) * (Nc + 1);
    for (int r = lane; r < (Nc + 1) * (Nc + 1); r += kWave) in[r] = src[r];
    __syncthreads();
    double th, ch;
    child_dir(c, 1, th, ch);
    const double *H = Hud + ((c >> 2) & 1 ? nrot : 0);
    rotate_wave(in, t1, tmp, ep, ec, Nc, H, -0.5 * kPi, ch, false);
    coax_wave(t1, Nc, t2, Np, SS, false);
    rotate_wave(t2, acc, tmp, ep, ec, Np, H, ch, -0.5 * kPi, true);
  }
  double2 *out = mps_p + (size_t)p * (Np + 1) * (Np + 1);
  for (int r = lane; r < (Np + 1) * (Np + 1); r += kWave) out[r] = acc[r];
}

// ---------------------------------------------------------------------
// L2L (src/pass.c:550-576, src/shift.c:403-585): parent local expansion
// to each flagged child (parent -> child direction)
// ---------------------------------------------------------------------
__global__ __launch_bounds__(64) void k_hel_l2l(
    const uint32_t *__restrict__ parents, unsigned npar,
    const uint32_t *__restrict__ cflag, int Np, int Nc,
    const double2 *__restrict__ mpr_p, double2 *__restrict__ mpr_c,
    const double *__restrict__ Hud, size_t nrot, const double2 *__restrict__ SS) {
  extern __shared__ double2 smem[];
  const int L = Nc > Np ? Nc : Np, L1 = (L + 1) * (L + 1);
  double2 *in = smem, *t1 = in + L1, *t2 = t1 + L1, *tmp = t2 + L1, *out = tmp + L1;
  double2 *ep = out + L1, *ec = ep + L + 1;
  const int lane = threadIdx.x;
  if (blockIdx.x >= npar) return;
  const uint32_t p = parents[blockIdx.x];
  const double2 *src = mpr_p + (size_t)p * (Np + 1) * (Np + 1);
  for (int r = lane; r < (Np + 1) * (Np + 1); r += kWave) in[r] = src[r];
  __syncthreads();
  for (int c = 0; c < 8; c++) {
    const uint32_t cb = 8 * p + c;
    if (cflag && !cflag[cb]) continue;
    double th, ch;
    child_dir(c, -1, th, ch);
    // parent -> child: e = -(0.5 - c bits); the z component flips the polar
    // angle, so the table index is that of the opposite z bit
    const double *H = Hud + (((c >> 2) & 1) ? 0 : nrot);
    rotate_wave(in, t1, tmp, ep, ec, Np, H, -0.5 * kPi, ch, false);
    coax_wave(t1, Np, t2, Nc, SS, false);
    rotate_wave(t2, out, tmp, ep, ec, Nc, H, ch, -0.5 * kPi, false);
    double2 *dst = mpr_c + (size_t)cb * (Nc + 1) * (Nc + 1);
    for (int r = lane; r < (Nc + 1) * (Nc + 1); r += kWave) dst[r] = cadd(dst[r], out[r]);
    __syncthreads();
  }
}

// ---------------------------------------------------------------------
// M2L (src/pass.c:56-105, :448-537): one wave per target box, its list-4
// sources in turn.  Offset geometry per slot o (dx = source - target):
// theta index, azimuth chi and length index; SR: one (E|F) table per
// length.
// ---------------------------------------------------------------------
__global__ __launch_bounds__(64) void k_hel_m2l(
    const uint32_t *__restrict__ targets, unsigned ntgt,
    const uint32_t *__restrict__ sstart, const uint32_t *__restrict__ send,
    int level, int Ns, int Nr, const double2 *__restrict__ mps,
    double2 *__restrict__ mpr, const double *__restrict__ Hth, size_t nrot,
    const int *__restrict__ th_idx, const double *__restrict__ chi,
    const int *__restrict__ rix, const double2 *__restrict__ SR, size_t ne) {
  extern __shared__ double2 smem[];
  const int L = Ns > Nr ? Ns : Nr, L1 = (L + 1) * (L + 1);
  double2 *in = smem, *t1 = in + L1, *t2 = t1 + L1, *tmp = t2 + L1, *acc = tmp + L1;
  double2 *ep = acc + L1, *ec = ep + L + 1;
  const int lane = threadIdx.x;
  if (blockIdx.x >= ntgt) return;
  const uint32_t tb = targets[blockIdx.x];
  const int ti = compact3(tb), tj = compact3(tb >> 1), tk = compact3(tb >> 2);
  const int px = ti & 1, py = tj & 1, pz = tk & 1, nbox = 1 << level;
  for (int r = lane; r < (Nr + 1) * (Nr + 1); r += kWave) acc[r] = make_double2(0.0, 0.0);
  for (int dz = -2 - pz; dz <= 3 - pz; dz++)
    for (int dy = -2 - py; dy <= 3 - py; dy++)
      for (int dx = -2 - px; dx <= 3 - px; dx++) {
        if (dx >= -1 && dx <= 1 && dy >= -1 && dy <= 1 && dz >= -1 && dz <= 1) continue;
        const int si = ti + dx, sj = tj + dy, sk = tk + dz;
        if (si < 0 || si >= nbox || sj < 0 || sj >= nbox || sk < 0 || sk >= nbox) continue;
        const uint64_t sb = spread3(si) | (spread3(sj) << 1) | (spread3(sk) << 2);
        if (send[sb] <= sstart[sb]) continue;
        const int o = (dx + 3) * 49 + (dy + 3) * 7 + dz + 3;
        const double *H = Hth + (size_t)th_idx[o] * nrot;
        const double c = chi[o];
        const double2 *src = mps + sb * (Ns + 1) * (Ns + 1);
        for (int r = lane; r < (Ns + 1) * (Ns + 1); r += kWave) in[r] = src[r];
        __syncthreads();
        rotate_wave(in, t1, tmp, ep, ec, Ns, H, -0.5 * kPi, c, false);
        coax_wave(t1, Ns, t2, Nr, SR + (size_t)rix[o] * ne, false);
        rotate_wave(t2, acc, tmp, ep, ec, Nr, H, c, -0.5 * kPi, true);
      }
  double2 *out = mpr + (size_t)tb * (Nr + 1) * (Nr + 1);
  for (int r = lane; r < (Nr + 1) * (Nr + 1); r += kWave) out[r] = cadd(out[r], acc[r]);
}

// ---------------------------------------------------------------------
// L2P + near field (src/local.c:32-310, src/expansions.c:275-344,
// src/gradient.c:282-399), one thread per target.  With
// R_n^m = j_n(kr) Pbar_n^|m| e^{i m phi} and the G&D eq 3.7 derivatives
// re-indexed so that every R_n^m (n <= N+1) is used once:
//   dz f = k sum R_n^m [a_n^m L_{n+1}^m - a_{n-1}^m L_{n-1}^m]
//   D+ f = k sum R_n^m [b_n^{-m} L_{n-1}^{m-1} - b_{n+1}^{m-1} L_{n+1}^{m-1}]
//   D- f = k sum R_n^m [b_n^m L_{n-1}^{m+1} - b_{n+1}^{-m-1} L_{n+1}^{m+1}]
//   grad_x = (D+ + D-)/2, grad_y = -i (D+ - D-)/2.
// Near field: G = e^{ikr}/(4 pi i k r), grad_x G = G (ik - 1/r) (x-y)/r,
// dipole term d nrm.grad_x G; pairs with r <= 1e-6 skipped.  Normals are
// sorted rows of 4 doubles like the positions.
// f[tid*fstr + ...]: potential (2 doubles) then gradient (6) when both
// are asked, accumulated.
// ---------------------------------------------------------------------
template <int NMAX>
__global__ __launch_bounds__(128) void k_hel_field(
    const double *__restrict__ xt, const uint32_t *__restrict__ tid, unsigned n,
    const uint32_t *__restrict__ start, const uint32_t *__restrict__ end,
    const double *__restrict__ xyz, const double2 *__restrict__ q,
    const double *__restrict__ nrm, const double2 *__restrict__ dip,
    unsigned level, double x0, double y0, double z0, double D, int N, double k,
    const double2 *__restrict__ mpr, int field, double *__restrict__ f, int fstr,
    const RecTab *__restrict__ RT) {
  const unsigned i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double x = xt[4 * i], y = xt[4 * i + 1], z = xt[4 * i + 2];
  const uint64_t code = point_code(x, y, z, x0, y0, z0, D);
  const uint64_t b = code >> (3 * (20 - level));
  const double w = D / (double)(1u << level);
  const double xc = x0 + (compact3(b) + 0.5) * w, yc = y0 + (compact3(b >> 1) + 0.5) * w,
               zc = z0 + (compact3(b >> 2) + 0.5) * w;
  double2 pot = make_double2(0.0, 0.0), gz = pot, dpl = pot, dmi = pot;
  const double2 *Lc = mpr + b * (size_t)(N + 1) * (N + 1);
  auto Lg = [&](int nn, int mm) -> double2 {
    if (nn < 0 || nn > N || mm > nn || -mm > nn) return make_double2(0.0, 0.0);
    return Lc[hidx(nn, mm)];
  };
  {
    double r, C, S, cp, sp, jn[NMAX + 2];
    sph(x - xc, y - yc, z - zc, r, C, S, cp, sp);
    bessel_j(N + 1, k * r, jn);
    const double phi = atan2(sp, cp);
    double pmm = 1.0 / sqrt(4.0 * kPi);
    for (int m = 0; m <= N + 1; m++) {
      if (m > 0) pmm *= sqrt((2.0 * m + 1) / (2.0 * m)) * S;
      double sn, cs;
      sincos((double)m * phi, &sn, &cs);
      double p = pmm, pprev = 0.0;
      for (int nn = m; nn <= N + 1; nn++) {
        const double a = jn[nn] * p;
        for (int sg = 1; sg >= -1; sg -= 2) {
          if (m == 0 && sg < 0) break;
          const int mm = sg * m;
          const double2 R = make_double2(a * cs, sg * a * sn);
          if ((field & 1) && nn <= N) pot = cadd(pot, cmul(Lg(nn, mm), R));
          if (field & 2) {
            const double2 cz = cadd(cscale(Lg(nn + 1, mm), tab_a(RT, nn, mm)),
                                    cscale(Lg(nn - 1, mm), -tab_a(RT, nn - 1, mm)));
            const double2 cpz = cadd(cscale(Lg(nn - 1, mm - 1), tab_b(RT, nn, -mm)),
                                     cscale(Lg(nn + 1, mm - 1), -tab_b(RT, nn + 1, mm - 1)));
            const double2 cmz = cadd(cscale(Lg(nn - 1, mm + 1), tab_b(RT, nn, mm)),
                                     cscale(Lg(nn + 1, mm + 1), -tab_b(RT, nn + 1, -mm - 1)));
            gz = cadd(gz, cmul(cz, R));
            dpl = cadd(dpl, cmul(cpz, R));
            dmi = cadd(dmi, cmul(cmz, R));
          }
        }
        const double pn = tab_pbar_up(RT, nn, m, C, p, pprev);
        pprev = p;
        p = pn;
      }
    }
  }
  double2 gx = cscale(cadd(dpl, dmi), 0.5 * k);
  const double2 dd = cscale(cadd(dpl, cscale(dmi, -1.0)), 0.5 * k);
  double2 gy = make_double2(dd.y, -dd.x);
  gz = cscale(gz, k);
  // near field: the 27 neighbours of the target's box
  const int bi = compact3(b), bj = compact3(b >> 1), bk = compact3(b >> 2), nb = 1 << level;
  for (int a = -1; a <= 1; a++)
    for (int bb = -1; bb <= 1; bb++)
      for (int c = -1; c <= 1; c++) {
        const int si = bi + a, sj = bj + bb, sk = bk + c;
        if (si < 0 || si >= nb || sj < 0 || sj >= nb || sk < 0 || sk >= nb) continue;
        const uint64_t sb = spread3(si) | (spread3(sj) << 1) | (spread3(sk) << 2);
        for (uint32_t s = start[sb]; s < end[sb]; s++) {
          const double rx = x - xyz[4 * s], ry = y - xyz[4 * s + 1], rz = z - xyz[4 * s + 2];
          const double r = sqrt(rx * rx + ry * ry + rz * rz);
          if (r <= 1e-6) continue;
          double sn, cs;
          sincos(k * r, &sn, &cs);
          // G = e^{ikr} / (4 pi i k r) = (sin - i cos) / (4 pi k r)
          const double g0 = 1.0 / (4.0 * kPi * k * r);
          const double2 G = make_double2(sn * g0, -cs * g0);
          const double2 dG = cmul(G, make_double2(-1.0 / r, k));  // G (ik - 1/r)
          const double2 G2 = cmul(G, cadd(cmul(make_double2(-1.0 / r, k), make_double2(-1.0 / r, k)),
                                          make_double2(1.0 / (r * r), 0.0)));
          const double ux = rx / r, uy = ry / r, uz = rz / r;
          if (q) {
            const double2 qq = q[s];
            const double2 qG = cmul(qq, G), qd = cmul(qq, dG);
            pot = cadd(pot, qG);
            gx = cadd(gx, cscale(qd, ux));
            gy = cadd(gy, cscale(qd, uy));
            gz = cadd(gz, cscale(qd, uz));
          }
          if (dip) {
            const double2 dq = dip[s];
            const double nx = nrm[4 * s], ny = nrm[4 * s + 1], nz = nrm[4 * s + 2];
            const double nr = nx * ux + ny * uy + nz * uz;
            pot = cadd(pot, cscale(cmul(dq, dG), nr));
            const double2 t1 = cscale(cmul(dq, G2), nr), t2 = cscale(cmul(dq, dG), 1.0 / r);
            gx = cadd(gx, cadd(cscale(t1, ux), cscale(t2, nx - nr * ux)));
            gy = cadd(gy, cadd(cscale(t1, uy), cscale(t2, ny - nr * uy)));
            gz = cadd(gz, cadd(cscale(t1, uz), cscale(t2, nz - nr * uz)));
          }
        }
      }
  double *o = f + (size_t)tid[i] * fstr;
  int at = 0;
  if (field & 1) {
    o[0] += pot.x;
    o[1] += pot.y;
    at = 2;
  }
  if (field & 2) {
    o[at + 0] += gx.x;
    o[at + 1] += gx.y;
    o[at + 2] += gy.x;
    o[at + 3] += gy.y;
    o[at + 4] += gz.x;
    o[at + 5] += gz.y;
  }
}

// ---------------------------------------------------------------------
// L2P + near field with one wavefront per target box (the same sums as
// k_hel_field).  Lane i takes target t0 + i of the box (64 at a time);
// the sources of the 27 neighbour boxes pass through LDS 64 at a time and
// every lane runs the same source loop, so the loads are shared and the
// lanes do not diverge.  T = double: FP64 throughout.  T = float: the
// local expansion holds L_n^m s_n in FP32 (sc: the level's scales s_n,
// divided out in FP64), the regular functions and the L2P sums are FP64,
// and the near-field kernel runs in FP32 (offsets x - y and the dipole
// factor n.(x - y) formed in FP64, the 1e-6 cutoff tested in FP64) with
// compensated (Neumaier) sums.
// ---------------------------------------------------------------------
__device__ __forceinline__ void sincos_t(double a, double *s, double *c) { sincos(a, s, c); }
__device__ __forceinline__ void sincos_t(float a, float *s, float *c) { sincosf(a, s, c); }

template <typename T>
struct CSum {  // compensated complex accumulator
  T sx, sy, cx, cy;
  __device__ __forceinline__ void zero() { sx = sy = cx = cy = T(0); }
  __device__ __forceinline__ static void add1(T &s, T &c, T v) {
    const T t = s + v;
    c += (fabs(s) >= fabs(v)) ? (s - t) + v : (v - t) + s;
    s = t;
  }
  __device__ __forceinline__ void add(T a, T b) {
    add1(sx, cx, a);
    add1(sy, cy, b);
  }
  __device__ __forceinline__ double2 value() const {
    return make_double2((double)sx + (double)cx, (double)sy + (double)cy);
  }
};

template <typename T, int NMAX>
__global__ __launch_bounds__(64) void k_hel_field_box(
    const uint32_t *__restrict__ tboxes, unsigned nbox, const uint32_t *__restrict__ tstart,
    const uint32_t *__restrict__ tend, const double *__restrict__ xt,
    const uint32_t *__restrict__ tid, const uint32_t *__restrict__ start,
    const uint32_t *__restrict__ end, const double *__restrict__ xyz,
    const double2 *__restrict__ q, const double *__restrict__ nrm,
    const double2 *__restrict__ dip, unsigned level, double x0, double y0, double z0,
    double D, int N, double k, const typename MF<T>::t2 *__restrict__ mpr,
    const double *__restrict__ sc, int field, double *__restrict__ f, int fstr) {
  typedef typename MF<T>::t2 T2;
  __shared__ double sxs[64][3];
  __shared__ double2 sqs[64], sds[64];
  __shared__ double sns[64][3];
  __shared__ double sinv[NMAX + 2];
  const int lane = threadIdx.x;
  const unsigned bi = blockIdx.x;
  if (bi >= nbox) return;  // uniform over the block
  for (int n = lane; n <= N + 1; n += 64) sinv[n] = sc ? 1.0 / sc[n] : 1.0;
  const uint32_t b = tboxes[bi], t0 = tstart[b], t1 = tend[b];
  const double w = D / (double)(1u << level);
  const int ci = compact3(b), cj = compact3((uint64_t)b >> 1), ck = compact3((uint64_t)b >> 2);
  const double xc = x0 + (ci + 0.5) * w, yc = y0 + (cj + 0.5) * w, zc = z0 + (ck + 0.5) * w;
  const T2 *Lc = mpr + (size_t)b * (N + 1) * (N + 1);
  const int nb = 1 << level;
  __syncthreads();
  for (uint32_t tb0 = t0; tb0 < t1; tb0 += 64) {
    const uint32_t it = tb0 + lane;
    const bool act = it < t1;
    const uint32_t ii = act ? it : t0;
    const double x = xt[4 * (size_t)ii], y = xt[4 * (size_t)ii + 1], z = xt[4 * (size_t)ii + 2];
    double2 pot = make_double2(0.0, 0.0), gz = pot, dpl = pot, dmi = pot;
    auto Lg = [&](int nn, int mm) -> double2 {
      if (nn < 0 || nn > N || mm > nn || -mm > nn) return make_double2(0.0, 0.0);
      const T2 v = Lc[hidx(nn, mm)];
      return make_double2((double)v.x * sinv[nn], (double)v.y * sinv[nn]);
    };
    {
      double r, C, S, cp, sp, jn[NMAX + 2];
      sph(x - xc, y - yc, z - zc, r, C, S, cp, sp);
      bessel_j(N + 1, k * r, jn);
      const double phi = atan2(sp, cp);
      double pmm = 1.0 / sqrt(4.0 * kPi);
      for (int m = 0; m <= N + 1; m++) {
        if (m > 0) pmm *= sqrt((2.0 * m + 1) / (2.0 * m)) * S;
        double sn, cs;
        sincos((double)m * phi, &sn, &cs);
        double p = pmm, pprev = 0.0;
        for (int nn = m; nn <= N + 1; nn++) {
          const double a = jn[nn] * p;
          for (int sg = 1; sg >= -1; sg -= 2) {
            if (m == 0 && sg < 0) break;
            const int mm = sg * m;
            const double2 R = make_double2(a * cs, sg * a * sn);
            if ((field & 1) && nn <= N) pot = cadd(pot, cmul(Lg(nn, mm), R));
            if (field & 2) {
              const double2 cz = cadd(cscale(Lg(nn + 1, mm), rec_a(nn, mm)),
                                      cscale(Lg(nn - 1, mm), -rec_a(nn - 1, mm)));
              const double2 cpz = cadd(cscale(Lg(nn - 1, mm - 1), rec_b(nn, -mm)),
                                       cscale(Lg(nn + 1, mm - 1), -rec_b(nn + 1, mm - 1)));
              const double2 cmz = cadd(cscale(Lg(nn - 1, mm + 1), rec_b(nn, mm)),
                                       cscale(Lg(nn + 1, mm + 1), -rec_b(nn + 1, -mm - 1)));
              gz = cadd(gz, cmul(cz, R));
              dpl = cadd(dpl, cmul(cpz, R));
              dmi = cadd(dmi, cmul(cmz, R));
            }
          }
          const double pn = pbar_up(nn, m, C, p, pprev);
          pprev = p;
          p = pn;
        }
      }
    }
    double2 gx = cscale(cadd(dpl, dmi), 0.5 * k);
    const double2 dd = cscale(cadd(dpl, cscale(dmi, -1.0)), 0.5 * k);
    double2 gy = make_double2(dd.y, -dd.x);
    gz = cscale(gz, k);
    // near field in precision T
    CSum<T> apot, agx, agy, agz;
    apot.zero();
    agx.zero();
    agy.zero();
    agz.zero();
    const T kk = (T)k, c4 = (T)(1.0 / (4.0 * kPi * k));
    for (int a = -1; a <= 1; a++)
      for (int bb = -1; bb <= 1; bb++)
        for (int c = -1; c <= 1; c++) {
          const int si = ci + a, sj = cj + bb, sk = ck + c;
          if (si < 0 || si >= nb || sj < 0 || sj >= nb || sk < 0 || sk >= nb) continue;
          const uint64_t sb = spread3(si) | (spread3(sj) << 1) | (spread3(sk) << 2);
          const uint32_t s0 = start[sb], s1 = end[sb];
          for (uint32_t sc0 = s0; sc0 < s1; sc0 += 64) {
            const uint32_t sl = sc0 + lane;
            __syncthreads();  // the previous tile's reads are done
            if (sl < s1) {
              sxs[lane][0] = xyz[4 * (size_t)sl];
              sxs[lane][1] = xyz[4 * (size_t)sl + 1];
              sxs[lane][2] = xyz[4 * (size_t)sl + 2];
              sqs[lane] = q ? q[sl] : make_double2(0.0, 0.0);
              if (dip) {
                sds[lane] = dip[sl];
                sns[lane][0] = nrm[4 * (size_t)sl];
                sns[lane][1] = nrm[4 * (size_t)sl + 1];
                sns[lane][2] = nrm[4 * (size_t)sl + 2];
              }
            }
            __syncthreads();
            const int cnt = (int)min(64u, s1 - sc0);
            for (int j = 0; j < cnt; j++) {
              const double rxd = x - sxs[j][0], ryd = y - sxs[j][1], rzd = z - sxs[j][2];
              const double r2 = rxd * rxd + ryd * ryd + rzd * rzd;
              if (r2 <= 1e-12) continue;
              const T rx = (T)rxd, ry = (T)ryd, rz = (T)rzd;
              const T r = sqrt((T)r2), ir = T(1) / r;
              T sn, cs;
              sincos_t(kk * r, &sn, &cs);
              // G = (sin - i cos) / (4 pi k r), G (ik - 1/r), G ((ik - 1/r)^2 + 1/r^2)
              const T g0 = c4 * ir;
              const T Gr = sn * g0, Gi = -cs * g0;
              const T dGr = -Gr * ir - Gi * kk, dGi = Gr * kk - Gi * ir;
              const T ux = rx * ir, uy = ry * ir, uz = rz * ir;
              if (q) {
                const T qr = (T)sqs[j].x, qi = (T)sqs[j].y;
                apot.add(qr * Gr - qi * Gi, qr * Gi + qi * Gr);
                const T qdr = qr * dGr - qi * dGi, qdi = qr * dGi + qi * dGr;
                agx.add(qdr * ux, qdi * ux);
                agy.add(qdr * uy, qdi * uy);
                agz.add(qdr * uz, qdi * uz);
              }
              if (dip) {
                const T dr = (T)sds[j].x, di = (T)sds[j].y;
                const T nx = (T)sns[j][0], ny = (T)sns[j][1], nz = (T)sns[j][2];
                // n.(x-y) in FP64: for close points on a smooth surface the
                // normal is nearly orthogonal to x - y, and the FP32 dot
                // product of unit vectors would cancel
                const T nr =
                    (T)(sns[j][0] * rxd + sns[j][1] * ryd + sns[j][2] * rzd) * ir;
                const T ddr = dr * dGr - di * dGi, ddi = dr * dGi + di * dGr;
                apot.add(ddr * nr, ddi * nr);
                // (ik - 1/r)^2 + 1/r^2 = 2/r^2 - k^2 - 2ik/r
                const T e2r = T(2) * ir * ir - kk * kk, e2i = -T(2) * kk * ir;
                const T g2r = Gr * e2r - Gi * e2i, g2i = Gr * e2i + Gi * e2r;
                const T t1r = (dr * g2r - di * g2i) * nr, t1i = (dr * g2i + di * g2r) * nr;
                const T t2r = ddr * ir, t2i = ddi * ir;
                agx.add(t1r * ux + t2r * (nx - nr * ux), t1i * ux + t2i * (nx - nr * ux));
                agy.add(t1r * uy + t2r * (ny - nr * uy), t1i * uy + t2i * (ny - nr * uy));
                agz.add(t1r * uz + t2r * (nz - nr * uz), t1i * uz + t2i * (nz - nr * uz));
              }
            }
          }
        }
    pot = cadd(pot, apot.value());
    gx = cadd(gx, agx.value());
    gy = cadd(gy, agy.value());
    gz = cadd(gz, agz.value());
    if (act) {
      double *o = f + (size_t)tid[it] * fstr;
      int at = 0;
      if (field & 1) {
        o[0] += pot.x;
        o[1] += pot.y;
        at = 2;
      }
      if (field & 2) {
        o[at + 0] += gx.x;
        o[at + 1] += gx.y;
        o[at + 2] += gy.x;
        o[at + 3] += gy.y;
        o[at + 4] += gz.x;
        o[at + 5] += gz.y;
      }
    }
  }
}

// ---------------------------------------------------------------------
// Mixed precision with one thread per target (the form of k_hel_field):
// the local expansion holds L_n^m s_n in FP32 (sc: the level's scales
// s_n, divided out in FP64 through 1/s_n in LDS), the regular functions
// and the L2P sums are FP64, and the near-field kernel runs in FP32 with
// compensated (Neumaier) sums, the offsets x - y, the 1e-6 cutoff and the
// dipole factor n.(x - y) in FP64.  Consecutive threads take targets of
// one box (they arrive sorted), so a wavefront's source loads coincide.
// On leaves of a few to a few tens of targets this keeps every lane busy,
// where one wavefront per target box leaves most of its lanes idle.
// ---------------------------------------------------------------------
template <int NMAX>
__global__ __launch_bounds__(128) void k_hel_field_mixed(
    const double *__restrict__ xt, const uint32_t *__restrict__ tid, unsigned n,
    const uint32_t *__restrict__ start, const uint32_t *__restrict__ end,
    const double *__restrict__ xyz, const double2 *__restrict__ q,
    const double *__restrict__ nrm, const double2 *__restrict__ dip,
    unsigned level, double x0, double y0, double z0, double D, int N, double k,
    const float2 *__restrict__ mpr, const double *__restrict__ sc, int field,
    double *__restrict__ f, int fstr, const RecTab *__restrict__ RT) {
  __shared__ double sinv[NMAX + 2];
  for (int t = threadIdx.x; t <= N + 1; t += blockDim.x) sinv[t] = 1.0 / sc[t];
  __syncthreads();
  const unsigned i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double x = xt[4 * i], y = xt[4 * i + 1], z = xt[4 * i + 2];
  const uint64_t code = point_code(x, y, z, x0, y0, z0, D);
  const uint64_t b = code >> (3 * (20 - level));
  const double w = D / (double)(1u << level);
  const double xc = x0 + (compact3(b) + 0.5) * w, yc = y0 + (compact3(b >> 1) + 0.5) * w,
               zc = z0 + (compact3(b >> 2) + 0.5) * w;
  double2 pot = make_double2(0.0, 0.0), gz = pot, dpl = pot, dmi = pot;
  const float2 *Lc = mpr + b * (size_t)(N + 1) * (N + 1);
  // branch-free (the bounds are the same in every lane): a coefficient
  // outside the expansion reads slot 0 with weight 0, so that the loads of
  // a step are issued together (9.0 -> 7.1 ms on the 1e6-point sphere; the
  // same form slowed the FP64 k_hel_field, 7.8 -> 8.5 ms)
  auto Lg = [&](int nn, int mm) -> double2 {
    const bool ok = !(nn < 0 || nn > N || mm > nn || -mm > nn);
    const float2 v = Lc[ok ? hidx(nn, mm) : 0];
    const double w = ok ? sinv[ok ? nn : 0] : 0.0;
    return make_double2((double)v.x * w, (double)v.y * w);
  };
  {
    double r, C, S, cp, sp, jn[NMAX + 2];
    sph(x - xc, y - yc, z - zc, r, C, S, cp, sp);
    bessel_j(N + 1, k * r, jn);
    const double phi = atan2(sp, cp);
    double pmm = 1.0 / sqrt(4.0 * kPi);
    for (int m = 0; m <= N + 1; m++) {
      if (m > 0) pmm *= sqrt((2.0 * m + 1) / (2.0 * m)) * S;
      double sn, cs;
      sincos((double)m * phi, &sn, &cs);
      double p = pmm, pprev = 0.0;
      for (int nn = m; nn <= N + 1; nn++) {
        const double a = jn[nn] * p;
        for (int sg = 1; sg >= -1; sg -= 2) {
          if (m == 0 && sg < 0) break;
          const int mm = sg * m;
          const double2 R = make_double2(a * cs, sg * a * sn);
          if ((field & 1) && nn <= N) pot = cadd(pot, cmul(Lg(nn, mm), R));
          if (field & 2) {
            const double2 cz = cadd(cscale(Lg(nn + 1, mm), tab_a(RT, nn, mm)),
                                    cscale(Lg(nn - 1, mm), -tab_a(RT, nn - 1, mm)));
            const double2 cpz = cadd(cscale(Lg(nn - 1, mm - 1), tab_b(RT, nn, -mm)),
                                     cscale(Lg(nn + 1, mm - 1), -tab_b(RT, nn + 1, mm - 1)));
            const double2 cmz = cadd(cscale(Lg(nn - 1, mm + 1), tab_b(RT, nn, mm)),
                                     cscale(Lg(nn + 1, mm + 1), -tab_b(RT, nn + 1, -mm - 1)));
            gz = cadd(gz, cmul(cz, R));
            dpl = cadd(dpl, cmul(cpz, R));
            dmi = cadd(dmi, cmul(cmz, R));
          }
        }
        const double pn = tab_pbar_up(RT, nn, m, C, p, pprev);
        pprev = p;
        p = pn;
      }
    }
  }
  double2 gx = cscale(cadd(dpl, dmi), 0.5 * k);
  const double2 dd = cscale(cadd(dpl, cscale(dmi, -1.0)), 0.5 * k);
  double2 gy = make_double2(dd.y, -dd.x);
  gz = cscale(gz, k);
  // near field in FP32, compensated sums
  CSum<float> apot, agx, agy, agz;
  apot.zero();
  agx.zero();
  agy.zero();
  agz.zero();
  const float kk = (float)k, c4 = (float)(1.0 / (4.0 * kPi * k));
  const int bi = compact3(b), bj = compact3(b >> 1), bk = compact3(b >> 2), nb = 1 << level;
  for (int a = -1; a <= 1; a++)
    for (int bb = -1; bb <= 1; bb++)
      for (int c = -1; c <= 1; c++) {
        const int si = bi + a, sj = bj + bb, sk = bk + c;
        if (si < 0 || si >= nb || sj < 0 || sj >= nb || sk < 0 || sk >= nb) continue;
        const uint64_t sb = spread3(si) | (spread3(sj) << 1) | (spread3(sk) << 2);
        for (uint32_t s = start[sb]; s < end[sb]; s++) {
          const double rxd = x - xyz[4 * s], ryd = y - xyz[4 * s + 1], rzd = z - xyz[4 * s + 2];
          const double r2 = rxd * rxd + ryd * ryd + rzd * rzd;
          if (r2 <= 1e-12) continue;
          const float rx = (float)rxd, ry = (float)ryd, rz = (float)rzd;
          const float r = sqrtf((float)r2), ir = 1.0f / r;
          float sn, cs;
          sincosf(kk * r, &sn, &cs);
          // G = (sin - i cos) / (4 pi k r), G (ik - 1/r), G ((ik - 1/r)^2 + 1/r^2)
          const float g0 = c4 * ir;
          const float Gr = sn * g0, Gi = -cs * g0;
          const float dGr = -Gr * ir - Gi * kk, dGi = Gr * kk - Gi * ir;
          const float ux = rx * ir, uy = ry * ir, uz = rz * ir;
          if (q) {
            const float qr = (float)q[s].x, qi = (float)q[s].y;
            apot.add(qr * Gr - qi * Gi, qr * Gi + qi * Gr);
            const float qdr = qr * dGr - qi * dGi, qdi = qr * dGi + qi * dGr;
            agx.add(qdr * ux, qdi * ux);
            agy.add(qdr * uy, qdi * uy);
            agz.add(qdr * uz, qdi * uz);
          }
          if (dip) {
            const float dr = (float)dip[s].x, di = (float)dip[s].y;
            const double nxd = nrm[4 * s], nyd = nrm[4 * s + 1], nzd = nrm[4 * s + 2];
            const float nx = (float)nxd, ny = (float)nyd, nz = (float)nzd;
            // n.(x-y) in FP64 (close points on a surface: n nearly orthogonal)
            const float nr = (float)(nxd * rxd + nyd * ryd + nzd * rzd) * ir;
            const float ddr = dr * dGr - di * dGi, ddi = dr * dGi + di * dGr;
            apot.add(ddr * nr, ddi * nr);
            // (ik - 1/r)^2 + 1/r^2 = 2/r^2 - k^2 - 2ik/r
            const float e2r = 2.0f * ir * ir - kk * kk, e2i = -2.0f * kk * ir;
            const float g2r = Gr * e2r - Gi * e2i, g2i = Gr * e2i + Gi * e2r;
            const float t1r = (dr * g2r - di * g2i) * nr, t1i = (dr * g2i + di * g2r) * nr;
            const float t2r = ddr * ir, t2i = ddi * ir;
            agx.add(t1r * ux + t2r * (nx - nr * ux), t1i * ux + t2i * (nx - nr * ux));
            agy.add(t1r * uy + t2r * (ny - nr * uy), t1i * uy + t2i * (ny - nr * uy));
            agz.add(t1r * uz + t2r * (nz - nr * uz), t1i * uz + t2i * (nz - nr * uz));
          }
        }
      }
  pot = cadd(pot, apot.value());
  gx = cadd(gx, agx.value());
  gy = cadd(gy, agy.value());
  gz = cadd(gz, agz.value());
  double *o = f + (size_t)tid[i] * fstr;
  int at = 0;
  if (field & 1) {
    o[0] += pot.x;
    o[1] += pot.y;
    at = 2;
  }
  if (field & 2) {
    o[at + 0] += gx.x;
    o[at + 1] += gx.y;
    o[at + 2] += gy.x;
    o[at + 3] += gy.y;
    o[at + 4] += gz.x;
    o[at + 5] += gz.y;
  }
}

// out (FP32, box-major, complex) = in (FP64) with degree n multiplied by
// sc[n], for the listed boxes; and the reverse (FP32 -> FP64, degree n
// multiplied by sc[n]: pass 1/s to unscale)
__global__ void k_hel_to_f32(const uint32_t *__restrict__ boxes, unsigned nbox, int N,
                             const double2 *__restrict__ in, const double *__restrict__ sc,
                             int add, float2 *__restrict__ out) {
  const int C = (N + 1) * (N + 1);
  const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (size_t)nbox * C) return;
  const size_t bx = e / C;
  const int c = (int)(e % C);
  const size_t at = (size_t)boxes[bx] * C + c;
  const double s = sc[isqrt_deg(c)];
  if (add)
    out[at] = make_float2((float)((double)out[at].x + in[at].x * s),
                          (float)((double)out[at].y + in[at].y * s));
  else
    out[at] = make_float2((float)(in[at].x * s), (float)(in[at].y * s));
}

__global__ void k_hel_to_f64(const uint32_t *__restrict__ boxes, unsigned nbox, int N,
                             const float2 *__restrict__ in, const double *__restrict__ sc,
                             double2 *__restrict__ out) {
  const int C = (N + 1) * (N + 1);
  const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (size_t)nbox * C) return;
  const size_t bx = e / C;
  const int c = (int)(e % C);
  const size_t at = (size_t)boxes[bx] * C + c;
  const double s = sc[isqrt_deg(c)];
  out[at] = make_double2((double)in[at].x * s, (double)in[at].y * s);
}

// ---------------------------------------------------------------------
// FP64 direct sum (src/direct.c:35-294): thread per target, sources
// streamed through LDS; same output layout as k_hel_field
// ---------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_hel_direct(
    const double *__restrict__ xyz, const double2 *__restrict__ q,
    const double *__restrict__ nrm, const double2 *__restrict__ dip,
    unsigned nsrc, const double *__restrict__ xt, unsigned ntgt, double k,
    int field, double cutoff, double *__restrict__ f, int fstr) {
  __shared__ double sx[256][4];
  __shared__ double2 sq[256], sd[256];
  __shared__ double sn3[256][3];
  const unsigned i = blockIdx.x * blockDim.x + threadIdx.x;
  const bool act = i < ntgt;
  const double x = act ? xt[4 * i] : 0.0, y = act ? xt[4 * i + 1] : 0.0,
               z = act ? xt[4 * i + 2] : 0.0;
  double2 pot = make_double2(0.0, 0.0), gx = pot, gy = pot, gz = pot;
  for (unsigned s0 = 0; s0 < nsrc; s0 += 256) {
    const unsigned s = s0 + threadIdx.x;
    __syncthreads();
    if (s < nsrc) {
      sx[threadIdx.x][0] = xyz[4 * s];
      sx[threadIdx.x][1] = xyz[4 * s + 1];
      sx[threadIdx.x][2] = xyz[4 * s + 2];
      sq[threadIdx.x] = q ? q[s] : make_double2(0.0, 0.0);
      sd[threadIdx.x] = dip ? dip[s] : make_double2(0.0, 0.0);
      sn3[threadIdx.x][0] = dip ? nrm[4 * s] : 0.0;
      sn3[threadIdx.x][1] = dip ? nrm[4 * s + 1] : 0.0;
      sn3[threadIdx.x][2] = dip ? nrm[4 * s + 2] : 0.0;
    }
    __syncthreads();
    const unsigned ns = nsrc - s0 < 256 ? nsrc - s0 : 256;
    for (unsigned j = 0; j < ns; j++) {
      const double rx = x - sx[j][0], ry = y - sx[j][1], rz = z - sx[j][2];
      const double r = sqrt(rx * rx + ry * ry + rz * rz);
      if (r <= cutoff) continue;
      double sn, cs;
      sincos(k * r, &sn, &cs);
      const double g0 = 1.0 / (4.0 * kPi * k * r);
      const double2 G = make_double2(sn * g0, -cs * g0);
      const double2 dG = cmul(G, make_double2(-1.0 / r, k));
      const double ux = rx / r, uy = ry / r, uz = rz / r;
      const double2 qG = cmul(sq[j], G), qd = cmul(sq[j], dG);
      pot = cadd(pot, qG);
      gx = cadd(gx, cscale(qd, ux));
      gy = cadd(gy, cscale(qd, uy));
      gz = cadd(gz, cscale(qd, uz));
      if (dip) {
        const double2 G2 = cmul(G, cadd(cmul(make_double2(-1.0 / r, k), make_double2(-1.0 / r, k)),
                                        make_double2(1.0 / (r * r), 0.0)));
        const double nx = sn3[j][0], ny = sn3[j][1], nz = sn3[j][2];
        const double nr = nx * ux + ny * uy + nz * uz;
        pot = cadd(pot, cscale(cmul(sd[j], dG), nr));
        const double2 t1 = cscale(cmul(sd[j], G2), nr), t2 = cscale(cmul(sd[j], dG), 1.0 / r);
        gx = cadd(gx, cadd(cscale(t1, ux), cscale(t2, nx - nr * ux)));
        gy = cadd(gy, cadd(cscale(t1, uy), cscale(t2, ny - nr * uy)));
        gz = cadd(gz, cadd(cscale(t1, uz), cscale(t2, nz - nr * uz)));
      }
    }
  }
  if (!act) return;
  double *o = f + (size_t)i * fstr;
  int at = 0;
  if (field & 1) {
    o[0] += pot.x;
    o[1] += pot.y;
    at = 2;
  }
  if (field & 2) {
    o[at + 0] += gx.x;
    o[at + 1] += gx.y;
    o[at + 2] += gy.x;
    o[at + 3] += gy.y;
    o[at + 4] += gz.x;
    o[at + 5] += gz.y;
  }
}

}  // namespace wbfmm

using namespace wbfmm;

// the process's recurrence table (built once, on the first call, on the
// calling stream; other streams read it once it is published)
static RecTab *g_rectab = nullptr;
static std::mutex g_rectab_mu;

static int rec_tab(const RecTab **out) {
  std::lock_guard<std::mutex> lk(g_rectab_mu);
  if (!g_rectab) {
    RecTab *t = nullptr;
    WBFMM_CHECK(hipMalloc((void **)&t, sizeof(RecTab)));
    WBFMM_LAUNCH("rec_tab", k_rec_tab, dim3((kRT * (2 * kRT + 1) + 255) / 256), dim3(256), 0, t);
    WBFMM_CHECK(hipStreamSynchronize(stream()));
    g_rectab = t;
  }
  *out = g_rectab;
  return 0;
}

extern "C" {

static size_t lds_vectors(int L, int nvec) {
  return sizeof(double2) * ((size_t)nvec * (L + 1) * (L + 1) + 2 * (L + 1));
}

int wbfmm_hip_hel_p2m(const uint32_t *boxes, unsigned nbox, const uint32_t *start,
                      const uint32_t *end, const double *xyz, const double *q,
                      const double *nrm, const double *dip, unsigned level,
                      const double *x0, double D, int N, double k, double *mps) {
  if (N < 0 || N > WBFMM_HIP_MAX_ORDER) return WBFMM_HIP_ERR_ORDER;
  if (nbox == 0) return 0;
  // points per batch: as many as 60 KiB of LDS hold, at most 8
  const size_t per = sizeof(double2) * (size_t)(N + 2) * (N + 3) / 2 + sizeof(double) * (N + 2);
  int pb = (int)((60 * 1024) / per);
  pb = pb < 1 ? 1 : (pb > kP2MBatch ? kP2MBatch : pb);
  const RecTab *RT = nullptr;
  {
    const int rc = rec_tab(&RT);
    if (rc) return rc;
  }
  {
    const int nc = ((N + 1) * (N + 1) + kP2MThreads - 1) / kP2MThreads;
#define HP2M(NCV) \
    WBFMM_LAUNCH("hel_p2m", (k_hel_p2m<double2, NCV>), dim3(nbox), dim3(kP2MThreads), per * pb, boxes, \
                 nbox, start, end, xyz, (const double2 *)q, nrm, (const double2 *)dip, level, \
                 x0[0], x0[1], x0[2], D, N, k, pb, (const double *)nullptr, (double2 *)mps, RT)
    switch (nc) {
      case 1: HP2M(1); break;
      case 2: HP2M(2); break;
      case 3: HP2M(3); break;
      case 4: HP2M(4); break;
      case 5: HP2M(5); break;
      case 6: HP2M(6); break;
      default: HP2M(kP2MCoef); break;
    }
#undef HP2M
  }
  return 0;
}

int wbfmm_hip_hel_p2m_f32(const uint32_t *boxes, unsigned nbox, const uint32_t *start,
                          const uint32_t *end, const double *xyz, const double *q,
                          const double *nrm, const double *dip, unsigned level,
                          const double *x0, double D, int N, double k, const double *scale,
                          float *mps) {
  if (N < 0 || N > WBFMM_HIP_MAX_ORDER) return WBFMM_HIP_ERR_ORDER;
  if (nbox == 0) return 0;
  const size_t per = sizeof(double2) * (size_t)(N + 2) * (N + 3) / 2 + sizeof(double) * (N + 2);
  int pb = (int)((60 * 1024) / per);
  pb = pb < 1 ? 1 : (pb > kP2MBatch ? kP2MBatch : pb);
  const RecTab *RT = nullptr;
  {
    const int rc = rec_tab(&RT);
    if (rc) return rc;
  }
  {
    const int nc = ((N + 1) * (N + 1) + kP2MThreads - 1) / kP2MThreads;
#define HP2M(NCV) \
    WBFMM_LAUNCH("hel_p2m_f32", (k_hel_p2m<float2, NCV>), dim3(nbox), dim3(kP2MThreads), per * pb, boxes, \
                 nbox, start, end, xyz, (const double2 *)q, nrm, (const double2 *)dip, level, \
                 x0[0], x0[1], x0[2], D, N, k, pb, scale, (float2 *)mps, RT)
    switch (nc) {
      case 1: HP2M(1); break;
      case 2: HP2M(2); break;
      case 3: HP2M(3); break;
      case 4: HP2M(4); break;
      case 5: HP2M(5); break;
      case 6: HP2M(6); break;
      default: HP2M(kP2MCoef); break;
    }
#undef HP2M
  }
  return 0;
}

int wbfmm_hip_hel_m2m(const uint32_t *parents, unsigned npar, const uint32_t *cstart,
                      const uint32_t *cend, int Nc, int Np, const double *mps_child,
                      double *mps_parent, const double *Hud, size_t nrot,
                      const double *SS) {
  if (npar == 0) return 0;
  const int L = Nc > Np ? Nc : Np;
  WBFMM_LAUNCH("hel_m2m", k_hel_m2m, dim3(npar), dim3(64), lds_vectors(L, 5), parents,
               npar, cstart, cend, Nc, Np, (const double2 *)mps_child,
               (double2 *)mps_parent, Hud, nrot, (const double2 *)SS);
  return 0;
}

int wbfmm_hip_hel_l2l(const uint32_t *parents, unsigned npar, const uint32_t *cflag,
                      int Np, int Nc, const double *mpr_parent, double *mpr_child,
                      const double *Hud, size_t nrot, const double *SS) {
  if (npar == 0) return 0;
  const int L = Nc > Np ? Nc : Np;
  WBFMM_LAUNCH("hel_l2l", k_hel_l2l, dim3(npar), dim3(64), lds_vectors(L, 5), parents,
               npar, cflag, Np, Nc, (const double2 *)mpr_parent, (double2 *)mpr_child,
               Hud, nrot, (const double2 *)SS);
  return 0;
}

int wbfmm_hip_hel_m2l(const uint32_t *targets, unsigned ntgt, const uint32_t *sstart,
                      const uint32_t *send, unsigned level, int Ns, int Nr,
                      const double *mps, double *mpr, const double *Hth, size_t nrot,
                      const int *th_idx, const double *chi, const int *rix,
                      const double *SR, size_t ne) {
  if (ntgt == 0) return 0;
  const int L = Ns > Nr ? Ns : Nr;
  WBFMM_LAUNCH("hel_m2l", k_hel_m2l, dim3(ntgt), dim3(64), lds_vectors(L, 5), targets,
               ntgt, sstart, send, (int)level, Ns, Nr, (const double2 *)mps,
               (double2 *)mpr, Hth, nrot, th_idx, chi, rix, (const double2 *)SR, ne);
  return 0;
}

int wbfmm_hip_hel_field_box(const uint32_t *tboxes, unsigned nbox, const uint32_t *tstart,
                            const uint32_t *tend, const double *xt, const uint32_t *tid,
                            const uint32_t *start, const uint32_t *end, const double *xyz,
                            const double *q, const double *nrm, const double *dip,
                            unsigned level, const double *x0, double D, int N, double k,
                            int prec, const void *mpr, const double *scale, int field,
                            double *f, int fstr) {
  if (nbox == 0) return 0;
  if (N < 0 || N > WBFMM_HIP_MAX_ORDER) return WBFMM_HIP_ERR_ORDER;
  if (prec != 32 && prec != 64) return WBFMM_HIP_ERR_ARG;
  if (prec == 32 && !scale) return WBFMM_HIP_ERR_ARG;
#define HB(TT, NM)                                                                             \
  WBFMM_LAUNCH(prec == 32 ? "hel_l2p_p2p_f32" : "hel_l2p_p2p", (k_hel_field_box<TT, NM>),     \
               dim3(nbox), dim3(64), 0, tboxes, nbox, tstart, tend, xt, tid, start, end, xyz, \
               (const double2 *)q, nrm, (const double2 *)dip, level, x0[0], x0[1], x0[2], D, \
               N, k, (const typename MF<TT>::t2 *)mpr, scale, field, f, fstr)
  if (prec == 32) {
    if (N <= 16) HB(float, 16);
    else if (N <= 24) HB(float, 24);
    else if (N <= 32) HB(float, 32);
    else HB(float, WBFMM_HIP_MAX_ORDER);
  } else {
    if (N <= 16) HB(double, 16);
    else if (N <= 24) HB(double, 24);
    else if (N <= 32) HB(double, 32);
    else HB(double, WBFMM_HIP_MAX_ORDER);
  }
#undef HB
  return 0;
}

int wbfmm_hip_hel_convert(int to32, int add, const uint32_t *boxes, unsigned nbox, int N,
                          const void *in, const double *scale, void *out) {
  if (nbox == 0) return 0;
  if (N < 0 || N > WBFMM_HIP_MAX_ORDER || !scale) return WBFMM_HIP_ERR_ARG;
  const size_t ne = (size_t)nbox * (N + 1) * (N + 1);
  if (to32)
    WBFMM_LAUNCH("hel_to_f32", k_hel_to_f32, dim3((unsigned)((ne + 255) / 256)), dim3(256), 0,
                 boxes, nbox, N, (const double2 *)in, scale, add, (float2 *)out);
  else
    WBFMM_LAUNCH("hel_to_f64", k_hel_to_f64, dim3((unsigned)((ne + 255) / 256)), dim3(256), 0,
                 boxes, nbox, N, (const float2 *)in, scale, (double2 *)out);
  return 0;
}

int wbfmm_hip_hel_field(const double *xt, const uint32_t *tid, unsigned n,
                        const uint32_t *start, const uint32_t *end, const double *xyz,
                        const double *q, const double *nrm, const double *dip,
                        unsigned level, const double *x0, double D, int N, double k,
                        const double *mpr, int field, double *f, int fstr) {
  if (n == 0) return 0;
  if (N < 0) return WBFMM_HIP_ERR_ORDER;
  const RecTab *RT = nullptr;
  {
    const int rc = rec_tab(&RT);
    if (rc) return rc;
  }
#define HF(NM)                                                                          \
  WBFMM_LAUNCH("hel_l2p_p2p", k_hel_field<NM>, dim3((n + 127) / 128), dim3(128), 0, xt, \
               tid, n, start, end, xyz, (const double2 *)q, nrm, (const double2 *)dip,  \
               level, x0[0], x0[1], x0[2], D, N, k, (const double2 *)mpr, field, f, fstr, RT)
  if (N <= 16) HF(16);
  else if (N <= 24) HF(24);
  else if (N <= 32) HF(32);
  else if (N <= WBFMM_HIP_MAX_ORDER) HF(WBFMM_HIP_MAX_ORDER);
  else return WBFMM_HIP_ERR_ORDER;
#undef HF
  return 0;
}

int wbfmm_hip_hel_field_f32(const double *xt, const uint32_t *tid, unsigned n,
                            const uint32_t *start, const uint32_t *end, const double *xyz,
                            const double *q, const double *nrm, const double *dip,
                            unsigned level, const double *x0, double D, int N, double k,
                            const float *mpr, const double *scale, int field, double *f,
                            int fstr) {
  if (n == 0) return 0;
  if (N < 0 || !scale) return N < 0 ? WBFMM_HIP_ERR_ORDER : WBFMM_HIP_ERR_ARG;
  const RecTab *RT = nullptr;
  {
    const int rc = rec_tab(&RT);
    if (rc) return rc;
  }
#define HF(NM)                                                                              \
  WBFMM_LAUNCH("hel_l2p_p2p_f32", k_hel_field_mixed<NM>, dim3((n + 127) / 128), dim3(128), 0, \
               xt, tid, n, start, end, xyz, (const double2 *)q, nrm, (const double2 *)dip,   \
               level, x0[0], x0[1], x0[2], D, N, k, (const float2 *)mpr, scale, field, f,    \
               fstr, RT)
  if (N <= 16) HF(16);
  else if (N <= 24) HF(24);
  else if (N <= 32) HF(32);
  else if (N <= WBFMM_HIP_MAX_ORDER) HF(WBFMM_HIP_MAX_ORDER);
  else return WBFMM_HIP_ERR_ORDER;
#undef HF
  return 0;
}

int wbfmm_hip_hel_direct(const double *xyz, const double *q, const double *nrm,
                         const double *dip, unsigned nsrc, const double *xt,
                         unsigned ntgt, double k, int field, double cutoff, double *f,
                         int fstr) {
  if (ntgt == 0) return 0;
  WBFMM_LAUNCH("hel_direct", k_hel_direct, dim3((ntgt + 255) / 256), dim3(256), 0, xyz,
               (const double2 *)q, nrm, (const double2 *)dip, nsrc, xt, ntgt, k, field,
               cutoff, f, fstr);
  return 0;
}

}  // extern "C"
