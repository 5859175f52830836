/*
 * wbfmm_oracle.c -- CPU restatement of the wbfmm octree FMM path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker: pytest, the smoke
 * check and the CPU-baseline leg of bench.py call it; the product library
 * (libwbfmm.so) never links or calls it.
 *
 * Each function names the reference file:line whose algorithm it restates.
 * The code is written from the mathematics those functions implement
 * (Gumerov & Duraiswami rotation/coaxial recurrences, normalised Legendre
 * recursions); it is not a copy of the reference source.
 */

#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <stdio.h>
#include <time.h>

#include <quadmath.h>

#include "wbfmm_oracle.h"

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

#define CUTOFF 1e-6  /* src/wbfmm.h:36 */

static double now_s(void)
{
  struct timespec ts;
  clock_gettime(CLOCK_MONOTONIC, &ts);
  return ts.tv_sec + 1e-9 * ts.tv_nsec;
}

/* index of (n,m), 0<=m<=n, in a triangular Legendre table */
#define PIDX(n, m) ((n) * ((n) + 1) / 2 + (m))
/* Laplace real-packed index, src/wbfmm-private.h:76 */
#define LIDX(n, m) ((n) * (n) + 2 * (m) - 1)
/* Helmholtz complex index, src/wbfmm.h:242 */
#define HIDX(n, m) ((n) * ((n) + 1) + (m))
/* rotation table index (nu,m,n), src/wbfmm-private.h:73-74 */
#define RIDX(nu, m, n) ((n) * (4 * (n) - 1) * ((n) + 1) / 6 + ((nu) + (n)) * ((n) + 1) + (m))

/* ======================================================================
 * special functions
 * ====================================================================== */

/*
 * Normalised associated Legendre functions, Gumerov-Duraiswami sign
 * convention (no Condon-Shortley phase) and 1/sqrt(4 pi) scaling:
 *   Pbar_n^m = sqrt((2n+1)/(4pi) (n-m)!/(n+m)!) P_n^m(C)
 * Restates src/util.c:67-116 (three-term recursion in n, diagonal
 * recursion for m=n).  P[PIDX(n,m)] for 0<=m<=n<=N.
 */
void orc_legendre(int N, double C, double S, double *P)
{
  int n, m;
  const double s4pi = 1.0 / sqrt(4.0 * M_PI);
  P[0] = s4pi;
  if (N == 0) return;
  P[PIDX(1, 0)] = C * sqrt(3.0) * s4pi;
  P[PIDX(1, 1)] = S * sqrt(1.5) * s4pi;
  for (n = 1; n < N; n++) {
    /* diagonal */
    P[PIDX(n + 1, n + 1)] = sqrt((2.0 * n + 3) / (2.0 * n + 2)) * S * P[PIDX(n, n)];
    for (m = 0; m <= n; m++) {
      double pm1 = (m <= n - 1) ? P[PIDX(n - 1, m)] : 0.0;
      double v = C * sqrt(2.0 * n + 1) * P[PIDX(n, m)] -
                 sqrt((double)(n * n - m * m) / (2 * n - 1)) * pm1;
      P[PIDX(n + 1, m)] = v / sqrt((double)((n + 1) * (n + 1) - m * m) / (2 * n + 3));
    }
  }
}

/* series for j_n(x), used when x < n+1 (src/util.c:118-145) */
static double bessel_j_series(int n, double x)
{
  double sc = 1.0 / (2 * n + 1), term = 1.0, s = 1.0, x2 = x * x;
  int q;
  for (q = 0; q < n; q++) sc *= x / (2 * q + 1);
  for (q = 0; q < 64 && fabs(term) > 1e-15; q++) {
    term *= -x2 / ((2 * q + 2) * (2 * n + 2 * (q + 1) + 1));
    s += term;
  }
  return s * sc;
}

/*
 * spherical Bessel j_0..j_N: upward recursion while x > n+1, series
 * otherwise (src/util.c:147-175).
 */
void orc_bessel_j(int N, double x, double *j)
{
  int n;
  if (x > 0.0) {
    j[0] = sin(x) / x;
    if (N >= 1) j[1] = j[0] / x - cos(x) / x;
  } else {
    j[0] = 1.0;
    if (N >= 1) j[1] = 0.0;
  }
  for (n = 1; n < N; n++) {
    if (x > n + 1)
      j[n + 1] = j[n] * (2 * n + 1) / x - j[n - 1];
    else
      j[n + 1] = bessel_j_series(n + 1, x);
  }
}

/*
 * spherical Hankel h_n^(1)(x) = j_n + i y_n by upward recursion,
 * h[2n], h[2n+1] = re, im (src/util.c:177-202).
 */
void orc_bessel_h(int N, double x, double *h)
{
  int n;
  h[0] = sin(x) / x;
  h[1] = -cos(x) / x;
  if (N >= 1) {
    h[2] = h[0] / x + h[1];
    h[3] = h[1] / x - h[0];
  }
  for (n = 1; n < N; n++) {
    h[2 * (n + 1) + 0] = h[2 * n + 0] * (2 * n + 1) / x - h[2 * (n - 1) + 0];
    h[2 * (n + 1) + 1] = h[2 * n + 1] * (2 * n + 1) / x - h[2 * (n - 1) + 1];
  }
}

/*
 * Truncation number of Gumerov & Duraiswami (JASA 2009) as used by the
 * reference, src/util.c:474-503.
 */
int orc_truncation_number(double k, double D, int level, double tol)
{
  /* p = (p_lo^4 + p_hi^4)^(1/4): low-frequency bound for separation
     ratio 2 and the high-frequency estimate ka + (3 ln(1/tol))^(2/3)
     (ka)^(1/3) / 2, ka = k * circumradius of a box at this level */
  const double del = 2.0;
  double ka = 0.5 * sqrt(3.0) * k * D / (1 << level);
  double plo = 1.0 - log(tol * (1.0 - 1.0 / del) * sqrt(1.0 - 1.0 / del)) / log(del);
  double phi = cbrt(3.0 * log(1.0 / tol));
  phi = ka + 0.5 * phi * phi * cbrt(ka);
  return (int)ceil(sqrt(sqrt(plo * plo * plo * plo + phi * phi * phi * phi)));
}

/* ======================================================================
 * Morton indexing (src/indexing.c:68-163, src/boxes.c:324-339)
 * ====================================================================== */

static uint64_t spread3(uint32_t v)
{
  uint64_t x = v & 0x1fffff;
  x = (x | (x << 32)) & 0x1f00000000ffffULL;
  x = (x | (x << 16)) & 0x1f0000ff0000ffULL;
  x = (x | (x << 8)) & 0x100f00f00f00f00fULL;
  x = (x | (x << 4)) & 0x10c30c30c30c30c3ULL;
  x = (x | (x << 2)) & 0x1249249249249249ULL;
  return x;
}

static uint32_t compact3(uint64_t x)
{
  x &= 0x1249249249249249ULL;
  x = (x ^ (x >> 2)) & 0x10c30c30c30c30c3ULL;
  x = (x ^ (x >> 4)) & 0x100f00f00f00f00fULL;
  x = (x ^ (x >> 8)) & 0x1f0000ff0000ffULL;
  x = (x ^ (x >> 16)) & 0x1f00000000ffffULL;
  x = (x ^ (x >> 32)) & 0x1fffff;
  return (uint32_t)x;
}

uint64_t orc_morton_encode(uint32_t x, uint32_t y, uint32_t z)
{
  return spread3(x) | (spread3(y) << 1) | (spread3(z) << 2);
}

void orc_morton_decode(uint64_t m, uint32_t *x, uint32_t *y, uint32_t *z)
{
  *x = compact3(m);
  *y = compact3(m >> 1);
  *z = compact3(m >> 2);
}

/* src/boxes.c:324-339: 20-bit grid coordinates, truncation toward zero */
uint64_t orc_point_index(const double *x, const double *x0, double D)
{
  uint32_t i = (uint32_t)((x[0] - x0[0]) / D * (1u << 20));
  uint32_t j = (uint32_t)((x[1] - x0[1]) / D * (1u << 20));
  uint32_t k = (uint32_t)((x[2] - x0[2]) / D * (1u << 20));
  return orc_morton_encode(i, j, k);
}

/* 27-neighbourhood clipped to the domain, src/indexing.c:250-286 */
int orc_box_neighbours(int level, uint64_t idx, uint64_t *nb)
{
  uint32_t i, j, k;
  int nbox = 1 << level, n = 0, ii, jj, kk;
  orc_morton_decode(idx, &i, &j, &k);
  for (ii = (int)i - 1; ii <= (int)i + 1; ii++) {
    if (ii < 0 || ii >= nbox) continue;
    for (jj = (int)j - 1; jj <= (int)j + 1; jj++) {
      if (jj < 0 || jj >= nbox) continue;
      for (kk = (int)k - 1; kk <= (int)k + 1; kk++) {
        if (kk < 0 || kk >= nbox) continue;
        nb[n++] = orc_morton_encode(ii, jj, kk);
      }
    }
  }
  return n;
}

/*
 * interaction list 4: children of the parent's neighbours that are not
 * neighbours of the box (src/indexing.c:330-377).  off[3n..3n+2] gets the
 * source-minus-target offset in box widths.
 */
int orc_box_list4(int level, uint64_t idx, uint64_t *list, int *off)
{
  uint32_t i, j, k;
  int nbox = 1 << level, n = 0, ii, jj, kk;
  int i0, j0, k0;
  orc_morton_decode(idx, &i, &j, &k);
  i0 = (int)i - (int)(i % 2) - 2;
  j0 = (int)j - (int)(j % 2) - 2;
  k0 = (int)k - (int)(k % 2) - 2;
  for (ii = i0; ii <= i0 + 5; ii++) {
    if (ii < 0 || ii >= nbox) continue;
    for (jj = j0; jj <= j0 + 5; jj++) {
      if (jj < 0 || jj >= nbox) continue;
      for (kk = k0; kk <= k0 + 5; kk++) {
        if (kk < 0 || kk >= nbox) continue;
        if (abs(ii - (int)i) <= 1 && abs(jj - (int)j) <= 1 &&
            abs(kk - (int)k) <= 1)
          continue;
        list[n] = orc_morton_encode(ii, jj, kk);
        if (off) {
          off[3 * n + 0] = ii - (int)i;
          off[3 * n + 1] = jj - (int)j;
          off[3 * n + 2] = kk - (int)k;
        }
        n++;
      }
    }
  }
  return n;
}

/* ======================================================================
 * rotations
 * ====================================================================== */

int orc_rotation_size(int N) { return RIDX(N + 1, 0, N + 1); }

/* G&D recursion coefficients (src/translations.c:182-203), evaluated in
   quad precision: the recursion below loses about 0.6 decimal digits per
   order in FP64 (2e-10 at N=12, 7e-6 at N=20 against a 60-digit run), so
   the oracle runs the reference's recursion with a 113-bit mantissa */
typedef __float128 qreal;

static qreal rec_a(int n, int m)
{
  m = abs(m);
  if (n < m) return 0;
  return sqrtq((qreal)(n + 1 + m) * (n + 1 - m) / ((qreal)(2 * n + 1) * (2 * n + 3)));
}

static qreal rec_b(int n, int m)
{
  qreal v;
  if (n < abs(m)) return 0;
  v = sqrtq((qreal)(n - m - 1) * (n - m) / ((qreal)(2 * n - 1) * (2 * n + 1)));
  return (m >= 0) ? v : -v;
}

/*
 * Rotation coefficients H_n^{nu,m}(theta), 0<=m<=n, -n<=nu<=n, n<=N.
 * m=0 entries from the normalised Legendre functions (G&D 5.48), then the
 * recursion G&D 5.55 advancing m; restates src/rotations.c:131-207.
 * H must hold orc_rotation_size(N) doubles.
 */
void orc_H_rotation(double *H, int N, double th)
{
  int L = 2 * N, n, m, nu, i;
  qreal C = cosq((qreal)th), S = sinq((qreal)th);
  int nrot = RIDX(L + 1, 0, L + 1);
  qreal *h = (qreal *)calloc(nrot, sizeof(qreal));
  qreal *P = (qreal *)malloc(sizeof(qreal) * (L + 2) * (L + 3) / 2);
  const qreal s4pi = 1 / sqrtq(4 * M_PIq);
  P[0] = s4pi;
  if (L >= 1) {
    P[PIDX(1, 0)] = C * sqrtq((qreal)3) * s4pi;
    P[PIDX(1, 1)] = S * sqrtq((qreal)1.5Q) * s4pi;
  }
  for (n = 1; n < L; n++) {
    P[PIDX(n + 1, n + 1)] = sqrtq((qreal)(2 * n + 3) / (2 * n + 2)) * S * P[PIDX(n, n)];
    for (m = 0; m <= n; m++) {
      qreal pm1 = (m <= n - 1) ? P[PIDX(n - 1, m)] : 0;
      P[PIDX(n + 1, m)] = (C * sqrtq((qreal)(2 * n + 1)) * P[PIDX(n, m)] -
                           sqrtq((qreal)(n * n - m * m) / (2 * n - 1)) * pm1) /
                          sqrtq((qreal)((n + 1) * (n + 1) - m * m) / (2 * n + 3));
    }
  }
  for (n = 0; n <= L; n++) {
    for (nu = 0; nu <= n; nu++) {
      qreal v = P[PIDX(n, nu)] / sqrtq((qreal)(2 * n + 1)) * sqrtq(4 * M_PIq);
      h[RIDX(nu, 0, n)] = v;
      h[RIDX(-nu, 0, n)] = v;
    }
  }
  for (m = 0; m <= N; m++) {
    for (n = m + 2; n <= L - m; n++) {
      qreal b = rec_b(n, m);
      for (nu = -n + 1; nu <= n - 1; nu++) {
        qreal b1 = rec_b(n, -nu - 1), b2 = rec_b(n, nu - 1);
        qreal a3 = rec_a(n - 1, nu);
        h[RIDX(nu, m + 1, n - 1)] =
            ((b1 * (1 - C) * h[RIDX(nu + 1, m, n)] -
              b2 * (1 + C) * h[RIDX(nu - 1, m, n)]) / 2 -
             a3 * S * h[RIDX(nu, m, n)]) / b;
      }
    }
  }
  for (i = 0; i < orc_rotation_size(N); i++) H[i] = (double)h[i];
  free(h);
  free(P);
}

/*
 * Rotation angles for a translation along d = target - source, expressed
 * in box widths from the source-minus-target offset (dx,dy,dz).  The frame
 * is that of src/util.c:224-262 (z' along d, x' = (d_y,-d_x,0)/|.|) and the
 * angles follow src/rotations.c:99-128; for offsets off the z axis
 * phi = -pi/2 and chi = atan2(d_y,d_x), which is what the reference's
 * tabulated angles hold (src/shifts.c:23-138, src/shift.c:628-724).
 */
void orc_shift_angles(int dx, int dy, int dz, double *th, double *ph,
                      double *ch, double *r)
{
  double ex = -dx, ey = -dy, ez = -dz;
  double rr = sqrt(ex * ex + ey * ey + ez * ez);
  *r = rr;
  *th = acos(ez / rr);
  if (dx == 0 && dy == 0) {
    *ph = 0.0;
    *ch = 0.0;
  } else {
    *ph = -0.5 * M_PI;
    *ch = atan2(ey, ex);
  }
}

/* ======================================================================
 * Laplace operators
 * ====================================================================== */

static void sph(const double *x0, const double *x, double *r, double *C,
                double *S, double *cp, double *sp)
{
  double dx = x[0] - x0[0], dy = x[1] - x0[1], dz = x[2] - x0[2];
  double rr = sqrt(dx * dx + dy * dy + dz * dz), rho;
  *r = rr;
  if (rr == 0.0) {
    *C = 1.0; *S = 0.0; *cp = 1.0; *sp = 0.0;
    return;
  }
  *C = dz / rr;
  rho = sqrt(dx * dx + dy * dy);
  *S = rho / rr;
  if (rho == 0.0) { *cp = 1.0; *sp = 0.0; }
  else { *cp = dx / rho; *sp = dy / rho; }
}

/*
 * P2M, monopole: M_n^m += q r^n Pbar_n^m e^{-i m phi} / (2n+1)
 * (src/laplace.c:35-129).  M dense, real-packed.
 */
void orc_lap_p2m(int N, const double *x0, const double *xs, double q,
                 double *M)
{
  double r, C, S, cp, sp, rn = 1.0, cm, sm, t;
  double P[(64 + 1) * (64 + 2) / 2];
  int n, m;
  sph(x0, xs, &r, &C, &S, &cp, &sp);
  orc_legendre(N, C, S, P);
  for (n = 0; n <= N; n++) {
    M[n * n] += q * rn * P[PIDX(n, 0)] / (2 * n + 1);
    cm = 1.0; sm = 0.0;
    for (m = 1; m <= n; m++) {
      t = cm * cp - sm * sp; sm = sm * cp + cm * sp; cm = t;
      M[LIDX(n, m) + 0] += q * rn * P[PIDX(n, m)] * cm / (2 * n + 1);
      M[LIDX(n, m) + 1] -= q * rn * P[PIDX(n, m)] * sm / (2 * n + 1);
    }
    rn *= r;
  }
}

/*
 * P2M, normal dipole of strength d and unit normal nrm: the expansion of
 * d n.grad_y [1/(4 pi |x-y|)], i.e. the directional derivative of the
 * monopole expansion with respect to the source position (the reference
 * builds the same coefficients by its own recurrences,
 * src/laplace.c:188-423).  Computed here by differentiating the regular
 * solid harmonics analytically.
 */
void orc_lap_p2m_normal(int N, const double *x0, const double *xs,
                        const double *nrm, double d, double *M)
{
  /* conj(R_n^m(y)) derivative along nrm, using the identities
     dz R_n^m = a R_{n-1}^m, (dx + i dy) R_n^m = -b1 R_{n-1}^{m+1},
     (dx - i dy) R_n^m = b2 R_{n-1}^{m-1} (see orc_lap_l_eval). */
  double r, C, S, cp, sp, rn, cm, sm, t;
  double P[(64 + 1) * (64 + 2) / 2];
  double Rre[(64 + 1) * (64 + 2) / 2], Rim[(64 + 1) * (64 + 2) / 2];
  int n, m;
  sph(x0, xs, &r, &C, &S, &cp, &sp);
  orc_legendre(N, C, S, P);
  rn = 1.0;
  for (n = 0; n <= N; n++) {
    cm = 1.0; sm = 0.0;
    for (m = 0; m <= n; m++) {
      if (m > 0) { t = cm * cp - sm * sp; sm = sm * cp + cm * sp; cm = t; }
      Rre[PIDX(n, m)] = rn * P[PIDX(n, m)] * cm;
      Rim[PIDX(n, m)] = rn * P[PIDX(n, m)] * sm;
    }
    rn *= r;
  }
  for (n = 1; n <= N; n++) {
    double f = sqrt((2.0 * n + 1) / (2.0 * n - 1));
    for (m = 0; m <= n; m++) {
      /* derivative of R_n^m(y) = A(y) + iB(y) */
      double gzr = 0, gzi = 0, dpr = 0, dpi = 0, dmr = 0, dmi = 0;
      double a = f * sqrt((double)(n - m) * (n + m));
      double b1 = f * sqrt((double)(n - m) * (n - m - 1));
      double b2 = f * sqrt((double)(n + m) * (n + m - 1));
      double gxr, gxi, gyr, gyi, dr, di;
      if (m <= n - 1) { gzr = a * Rre[PIDX(n - 1, m)]; gzi = a * Rim[PIDX(n - 1, m)]; }
      if (m + 1 <= n - 1) { dpr = -b1 * Rre[PIDX(n - 1, m + 1)]; dpi = -b1 * Rim[PIDX(n - 1, m + 1)]; }
      if (m >= 1) {
        dmr = b2 * Rre[PIDX(n - 1, m - 1)]; dmi = b2 * Rim[PIDX(n - 1, m - 1)];
      } else {
        /* R_{n-1}^{-1} = -conj(R_{n-1}^1) scaled: use D- R_n^0 = conj(D+ R_n^0) */
        dmr = dpr; dmi = -dpi;
      }
      /* dx = (D+ + D-)/2, dy = (D+ - D-)/(2i) */
      gxr = 0.5 * (dpr + dmr); gxi = 0.5 * (dpi + dmi);
      gyr = 0.5 * (dpi - dmi); gyi = -0.5 * (dpr - dmr);
      dr = nrm[0] * gxr + nrm[1] * gyr + nrm[2] * gzr;
      di = nrm[0] * gxi + nrm[1] * gyi + nrm[2] * gzi;
      /* M_n^m += d conj(R_n^m) / (2n+1) */
      if (m == 0)
        M[n * n] += d * dr / (2 * n + 1);
      else {
        M[LIDX(n, m) + 0] += d * dr / (2 * n + 1);
        M[LIDX(n, m) + 1] -= d * di / (2 * n + 1);
      }
    }
  }
}

/*
 * singular-expansion evaluation, src/laplace.c:426-494:
 * f = sum r^{-n-1} [M_n^0 Pbar_n^0 + 2 sum_m Pbar_n^m Re(M_n^m e^{i m phi})]
 */
double orc_lap_m_eval(int N, const double *x0, const double *M,
                      const double *xf)
{
  double r, C, S, cp, sp, rn, cm, sm, t, f = 0.0;
  double P[(64 + 1) * (64 + 2) / 2];
  int n, m;
  sph(x0, xf, &r, &C, &S, &cp, &sp);
  orc_legendre(N, C, S, P);
  rn = 1.0 / r;
  for (n = 0; n <= N; n++) {
    f += M[n * n] * rn * P[PIDX(n, 0)];
    cm = 1.0; sm = 0.0;
    for (m = 1; m <= n; m++) {
      t = cm * cp - sm * sp; sm = sm * cp + cm * sp; cm = t;
      f += 2.0 * P[PIDX(n, m)] * rn *
           (M[LIDX(n, m)] * cm - M[LIDX(n, m) + 1] * sm);
    }
    rn /= r;
  }
  return f;
}

/*
 * local-expansion evaluation (L2P).  field&1: potential into f[0]
 * (src/local-laplace.c:333-398); field&2: gradient into f[1..3]
 * (src/local-laplace.c:525-630 with the derivative identities of
 * src/wbfmm-derivatives.h:255-314):
 *   dz R_n^m       =  a_n^m  R_{n-1}^m,     a  = sqrt((2n+1)/(2n-1)(n-m)(n+m))
 *   (dx+i dy) R_n^m = -b1_n^m R_{n-1}^{m+1}, b1 = sqrt((2n+1)/(2n-1)(n-m)(n-m-1))
 *   (dx-i dy) R_n^m =  b2_n^m R_{n-1}^{m-1}, b2 = sqrt((2n+1)/(2n-1)(n+m)(n+m-1))
 * f is accumulated into.
 */
void orc_lap_l_eval(int N, const double *x0, const double *L,
                    const double *xf, int field, double *f)
{
  double r, C, S, cp, sp, rn, cm, sm, t;
  double P[(64 + 1) * (64 + 2) / 2];
  double Rre[(64 + 1) * (64 + 2) / 2], Rim[(64 + 1) * (64 + 2) / 2];
  int n, m;
  sph(x0, xf, &r, &C, &S, &cp, &sp);
  orc_legendre(N, C, S, P);
  rn = 1.0;
  for (n = 0; n <= N; n++) {
    cm = 1.0; sm = 0.0;
    for (m = 0; m <= n; m++) {
      if (m > 0) { t = cm * cp - sm * sp; sm = sm * cp + cm * sp; cm = t; }
      Rre[PIDX(n, m)] = rn * P[PIDX(n, m)] * cm;
      Rim[PIDX(n, m)] = rn * P[PIDX(n, m)] * sm;
    }
    rn *= r;
  }
  if (field & 1) {
    double phi = 0.0;
    for (n = 0; n <= N; n++) {
      phi += L[n * n] * Rre[PIDX(n, 0)];
      for (m = 1; m <= n; m++)
        phi += 2.0 * (L[LIDX(n, m)] * Rre[PIDX(n, m)] -
                      L[LIDX(n, m) + 1] * Rim[PIDX(n, m)]);
    }
    f[0] += phi;
  }
  if (field & 2) {
    double gx = 0, gy = 0, gz = 0;
    for (n = 1; n <= N; n++) {
      double fc = sqrt((2.0 * n + 1) / (2.0 * n - 1));
      for (m = 0; m <= n; m++) {
        double lr, li, a, b1, b2, dpr = 0, dpi = 0;
        if (m == 0) { lr = L[n * n]; li = 0.0; }
        else { lr = L[LIDX(n, m)]; li = L[LIDX(n, m) + 1]; }
        a = fc * sqrt((double)(n - m) * (n + m));
        b1 = fc * sqrt((double)(n - m) * (n - m - 1));
        b2 = fc * sqrt((double)(n + m) * (n + m - 1));
        /* d/dz of L R_n^m (+ conj for m>0) */
        if (m <= n - 1) {
          double v = a * (lr * Rre[PIDX(n - 1, m)] - li * Rim[PIDX(n - 1, m)]);
          gz += (m == 0) ? v : 2.0 * v;
        }
        /* D+ of L R_n^m: -b1 L R_{n-1}^{m+1} */
        if (m + 1 <= n - 1) {
          double rr = Rre[PIDX(n - 1, m + 1)], ri = Rim[PIDX(n - 1, m + 1)];
          dpr += -b1 * (lr * rr - li * ri);
          dpi += -b1 * (lr * ri + li * rr);
        }
        /* D+ of conj(L) conj(R_n^m), m>=1: b2 conj(L) conj(R_{n-1}^{m-1}) */
        if (m >= 1) {
          double rr = Rre[PIDX(n - 1, m - 1)], ri = Rim[PIDX(n - 1, m - 1)];
          dpr += b2 * (lr * rr - li * ri);
          dpi += -b2 * (lr * ri + li * rr);
        }
        gx += dpr;
        gy += dpi;
      }
    }
    f[1] += gx;
    f[2] += gy;
    f[3] += gz;
  }
}

/*
 * Second derivatives of a Laplace local expansion at xf, added into
 * h[0..5] = (xx, yy, zz, xy, yz, zx): the quantities the reference's
 * curl-gradient evaluation combines (src/local-laplace.c:752-1011, with
 * the second-derivative macros of src/wbfmm-derivatives.h).  Restated
 * with ladder operators on S_n^mu = r^n Pbar_n^|mu| e^{i mu phi}
 * (S_n^-mu = conj S_n^mu, and L_n^-mu = conj L_n^mu, so that the field
 * is sum_{n,mu} L_n^mu S_n^mu):
 *   d/dz S_n^mu = a(n,|mu|) S_{n-1}^mu,   a(n,m) = sqrt(k_n (n-m)(n+m)),
 *   D+ S_n^mu = s(mu) b(n,mu) S_{n-1}^{mu+1},
 *               b(n,mu) = sqrt(k_n (n-mu)(n-mu-1)), s = -1 (mu >= 0), +1,
 * k_n = (2n+1)/(2n-1), D+ = d/dx + i d/dy (the identities the gradient
 * above uses).  With zz = d2/dz2, V = D+ d/dz and W = D+ D+ of the field:
 *   xx = (Re W - zz)/2, yy = -(Re W + zz)/2, xy = Im W / 2,
 *   zx = Re V, yz = Im V   (D+ D- = -d2/dz2 for a harmonic field).
 */
static double lad_a(int n, int m)
{
  return n >= 1 ? sqrt((2.0 * n + 1) / (2.0 * n - 1) * (double)(n - m) * (n + m)) : 0.0;
}

static double lad_b(int n, int mu)
{
  return n >= 1 ? sqrt((2.0 * n + 1) / (2.0 * n - 1) * (double)(n - mu) * (n - mu - 1)) : 0.0;
}

void orc_lap_l_hess(int N, const double *x0, const double *L, const double *xf, double *h)
{
  double r, C, S, cp, sp, rn, cm, sm, t;
  double P[(64 + 1) * (64 + 2) / 2];
  double Rre[(64 + 1) * (64 + 2) / 2], Rim[(64 + 1) * (64 + 2) / 2];
  double zz = 0.0, vr = 0.0, vi = 0.0, wr = 0.0, wi = 0.0;
  int n, m, mu;
  sph(x0, xf, &r, &C, &S, &cp, &sp);
  orc_legendre(N, C, S, P);
  rn = 1.0;
  for (n = 0; n <= N; n++) {
    cm = 1.0; sm = 0.0;
    for (m = 0; m <= n; m++) {
      if (m > 0) { t = cm * cp - sm * sp; sm = sm * cp + cm * sp; cm = t; }
      Rre[PIDX(n, m)] = rn * P[PIDX(n, m)] * cm;
      Rim[PIDX(n, m)] = rn * P[PIDX(n, m)] * sm;
    }
    rn *= r;
  }
  for (n = 2; n <= N; n++) {
    for (mu = -n; mu <= n; mu++) {
      const int am = mu < 0 ? -mu : mu;
      double lr, li;
      if (am == 0) { lr = L[n * n]; li = 0.0; }
      else { lr = L[LIDX(n, am)]; li = L[LIDX(n, am) + 1]; }
      if (mu < 0) li = -li;
      /* (c, nu): coefficient c of S_{n-2}^nu */
#define ADD_TERM(cre, cim, nu, ore, oim)                                          \
      do {                                                                         \
        const int an = (nu) < 0 ? -(nu) : (nu);                                   \
        if (an <= n - 2) {                                                         \
          double sr = Rre[PIDX(n - 2, an)], si = Rim[PIDX(n - 2, an)];             \
          if ((nu) < 0) si = -si;                                                  \
          ore += (cre) * sr - (cim) * si;                                          \
          oim += (cre) * si + (cim) * sr;                                          \
        }                                                                          \
      } while (0)
      {
        /* d2/dz2: a(n,|mu|) a(n-1,|mu|) */
        const double c = lad_a(n, am) * lad_a(n - 1, am);
        double dummy = 0.0;
        ADD_TERM(c * lr, c * li, mu, zz, dummy);
      }
      {
        /* D+ d/dz: a(n,|mu|) s(mu) b(n-1,mu) */
        const double c = lad_a(n, am) * (mu >= 0 ? -1.0 : 1.0) * lad_b(n - 1, mu);
        ADD_TERM(c * lr, c * li, mu + 1, vr, vi);
      }
      {
        /* D+ D+: s(mu) b(n,mu) s(mu+1) b(n-1,mu+1) */
        const double c = (mu >= 0 ? -1.0 : 1.0) * lad_b(n, mu) * (mu + 1 >= 0 ? -1.0 : 1.0) *
                         lad_b(n - 1, mu + 1);
        ADD_TERM(c * lr, c * li, mu + 2, wr, wi);
      }
#undef ADD_TERM
    }
  }
  h[0] += 0.5 * (wr - zz);
  h[1] += -0.5 * (wr + zz);
  h[2] += zz;
  h[3] += 0.5 * wi;
  h[4] += vi;
  h[5] += vr;
}

/*
 * closed-form Laplace coaxial coefficients (src/laplace.c:542-569,
 * :630-649, :698-725); kind 0 = SS, 1 = RR, 2 = SR.  The powers of the
 * translation distance are applied by orc_lap_coaxial.
 */
double orc_lap_coaxial_cfft(int kind, int n, int nu, int m)
{
  double c;
  if (kind == 0) {       /* SS, n >= nu */
    c = lgamma(n - m + 1) + lgamma(n + m + 1) - lgamma(nu - m + 1) -
        lgamma(nu + m + 1) - 2.0 * lgamma(n - nu + 1);
    c += log((2.0 * nu + 1) / (2.0 * n + 1));
    return exp(0.5 * c);
  }
  if (kind == 1) {       /* RR, nu >= n */
    c = lgamma(nu - m + 1) + lgamma(nu + m + 1) - lgamma(n - m + 1) -
        lgamma(n + m + 1) - 2.0 * lgamma(nu - n + 1);
    c += log((2.0 * nu + 1) / (2.0 * n + 1));
    return exp(0.5 * c);
  }
  c = 2.0 * lgamma(n + nu + 1) - lgamma(n - m + 1) - lgamma(n + m + 1) -
      lgamma(nu - m + 1) - lgamma(nu + m + 1);
  c += log((2.0 * nu + 1) / (2.0 * n + 1));
  c = exp(0.5 * c);
  return ((n + m) % 2) ? -c : c;
}

/* coefficient cache, the counterpart of the reference's global tables
   filled by wbfmm_laplace_coaxial_translate_init (src/laplace.c:727-792);
   cf[kind][(m*(LC+1)+n)*(LC+1)+nu] */
#define LC 64
static double *lap_cf[3];
static int lap_cf_N = -1;

static void lap_cf_init(int N)
{
  int kind, n, nu, m;
  if (N <= lap_cf_N) return;
  if (N > LC) N = LC;
  for (kind = 0; kind < 3; kind++) {
    free(lap_cf[kind]);
    lap_cf[kind] = calloc((size_t)(LC + 1) * (LC + 1) * (LC + 1), sizeof(double));
    for (m = 0; m <= N; m++)
      for (n = m; n <= N; n++)
        for (nu = m; nu <= N; nu++) {
          if (kind == 0 && nu > n) continue;
          if (kind == 1 && nu < n) continue;
          lap_cf[kind][((size_t)m * (LC + 1) + n) * (LC + 1) + nu] =
              orc_lap_coaxial_cfft(kind, n, nu, m);
        }
  }
  lap_cf_N = N;
}

/*
 * coaxial translation along z by t (src/laplace.c:496-696):
 *   SS: Co_n^m = sc Co + sum_{nu<=n} c (-t)^{n-nu} Ci_nu^m
 *   RR: Co_n^m = sc Co + sum_{nu>=n} c t^{nu-n} Ci_nu^m
 *   SR: Co_n^m = sc Co + sum_{nu} c / t^{n+nu+1} Ci_nu^m   (t signed)
 * Co and Ci dense real-packed.
 */
void orc_lap_coaxial(int kind, double *Co, int No, const double *Ci, int Ni,
                     double t, double sc)
{
  int n, m, nu, c;
  double tp[160];
  int nc = (No + 1) * (No + 1);
  double out[(LC + 1) * (LC + 1)];
  lap_cf_init(No > Ni ? No : Ni);
  memset(out, 0, sizeof(double) * nc);
  tp[0] = 1.0;
  for (n = 1; n < 160; n++) tp[n] = tp[n - 1] * t;
  for (m = 0; m <= No; m++) {
    for (n = m; n <= No; n++) {
      int lo, hi;
      const double *cf = lap_cf[kind] + ((size_t)m * (LC + 1) + n) * (LC + 1);
      double sr = 0.0, si = 0.0;
      if (kind == 0) { lo = m; hi = (n < Ni) ? n : Ni; }
      else if (kind == 1) { lo = (n > m) ? n : m; hi = Ni; }
      else { lo = m; hi = Ni; }
      for (nu = lo; nu <= hi; nu++) {
        double c1 = cf[nu];
        if (kind == 0) c1 *= ((n - nu) % 2 ? -1.0 : 1.0) * tp[n - nu];
        else if (kind == 1) c1 *= tp[nu - n];
        else c1 /= tp[n + nu + 1];
        if (m == 0)
          sr += c1 * Ci[nu * nu];
        else {
          sr += c1 * Ci[LIDX(nu, m) + 0];
          si += c1 * Ci[LIDX(nu, m) + 1];
        }
      }
      if (m == 0) out[n * n] = sr;
      else { out[LIDX(n, m)] = sr; out[LIDX(n, m) + 1] = si; }
    }
  }
  /* sc == 0 assigns (the output buffer may hold garbage) */
  for (c = 0; c < nc; c++) Co[c] = (sc == 0.0) ? out[c] : sc * Co[c] + out[c];
}

/*
 * rotation of a real-packed expansion, src/rotations-laplace.c:36-118:
 *   Co_n^nu = sc Co + e^{-i nu ph} sum_{m=-n}^{n} H_n^{nu,m} e^{i m ch} Ci_n^m
 * with Ci_n^{-m} = conj(Ci_n^m) and H_n^{nu,-m} = H_n^{-nu,m}.
 */
void orc_lap_rotate(double *Co, const double *Ci, int N, const double *H,
                    double ph, double ch, double sc)
{
  int n, m, nu;
  double cc[LC + 2], sc_[LC + 2], cp[LC + 2], sp[LC + 2];
  cc[0] = 1.0; sc_[0] = 0.0; cp[0] = 1.0; sp[0] = 0.0;
  for (m = 1; m <= N; m++) {
    cc[m] = cos(m * ch); sc_[m] = sin(m * ch);
    cp[m] = cos(m * ph); sp[m] = sin(m * ph);
  }
  for (n = 0; n <= N; n++) {
    double zr[LC + 1], zi[LC + 1];
    for (m = 1; m <= n; m++) {
      double cr = Ci[LIDX(n, m)], ci = Ci[LIDX(n, m) + 1];
      zr[m] = cr * cc[m] - ci * sc_[m];
      zi[m] = cr * sc_[m] + ci * cc[m];
    }
    for (nu = 0; nu <= n; nu++) {
      const double *hp = H + RIDX(nu, 0, n), *hm = H + RIDX(-nu, 0, n);
      double tr = hp[0] * Ci[n * n], ti = 0.0;
      for (m = 1; m <= n; m++) {
        tr += (hp[m] + hm[m]) * zr[m];
        ti += (hp[m] - hm[m]) * zi[m];
      }
      if (nu == 0) {
        Co[n * n] = (sc == 0.0) ? tr : sc * Co[n * n] + tr;
      } else {
        double vr = cp[nu] * tr + sp[nu] * ti, vi = cp[nu] * ti - sp[nu] * tr;
        if (sc == 0.0) {
          Co[LIDX(n, nu) + 0] = vr;
          Co[LIDX(n, nu) + 1] = vi;
        } else {
          Co[LIDX(n, nu) + 0] = sc * Co[LIDX(n, nu) + 0] + vr;
          Co[LIDX(n, nu) + 1] = sc * Co[LIDX(n, nu) + 1] + vi;
        }
      }
    }
  }
}

/* ======================================================================
 * FP64 direct sums
 * ====================================================================== */

/*
 * Laplace direct sum (src/direct-laplace.c:200-336).  field&1 -> f[4i],
 * field&2 -> f[4i+1..3].  Sources: monopoles q (may be NULL) and normal
 * dipoles d with unit normals nrm (may be NULL).  Pairs with r <= cutoff
 * are skipped (cutoff 0: none skipped, as in the reference direct sum).
 */
void orc_lap_direct(int nsrc, const double *xs, const double *q,
                    const double *nrm, const double *d,
                    int ntgt, const double *xt, int field, double cutoff,
                    double *f)
{
  int i, j;
  const double c4 = 0.25 / M_PI;
  const int fw = (field & 4) ? 10 : 4;
  for (i = 0; i < ntgt; i++) {
    double phi = 0, gx = 0, gy = 0, gz = 0, h[6] = {0, 0, 0, 0, 0, 0};
    for (j = 0; j < nsrc; j++) {
      double rx = xt[3 * i] - xs[3 * j], ry = xt[3 * i + 1] - xs[3 * j + 1];
      double rz = xt[3 * i + 2] - xs[3 * j + 2];
      double r2 = rx * rx + ry * ry + rz * rz, r, ir, ir3;
      if (r2 <= cutoff * cutoff) continue;
      r = sqrt(r2); ir = 1.0 / r; ir3 = ir * ir * ir;
      if (q) {
        phi += q[j] * ir;
        gx -= q[j] * rx * ir3; gy -= q[j] * ry * ir3; gz -= q[j] * rz * ir3;
        if (field & 4) {
          /* second derivatives of q/r: q (3 r_a r_b / r^5 - delta_ab / r^3) */
          const double ir5 = ir3 * ir * ir;
          h[0] += q[j] * (3.0 * rx * rx * ir5 - ir3);
          h[1] += q[j] * (3.0 * ry * ry * ir5 - ir3);
          h[2] += q[j] * (3.0 * rz * rz * ir5 - ir3);
          h[3] += q[j] * 3.0 * rx * ry * ir5;
          h[4] += q[j] * 3.0 * ry * rz * ir5;
          h[5] += q[j] * 3.0 * rz * rx * ir5;
        }
      }
      if (d) {
        double nr = nrm[3 * j] * rx + nrm[3 * j + 1] * ry + nrm[3 * j + 2] * rz;
        double ir5 = ir3 * ir * ir;
        phi += d[j] * nr * ir3;
        gx += d[j] * (nrm[3 * j] * ir3 - 3.0 * nr * rx * ir5);
        gy += d[j] * (nrm[3 * j + 1] * ir3 - 3.0 * nr * ry * ir5);
        gz += d[j] * (nrm[3 * j + 2] * ir3 - 3.0 * nr * rz * ir5);
      }
    }
    if (field & 1) f[fw * i] += phi * c4;
    if (field & 2) {
      f[fw * i + 1] += gx * c4; f[fw * i + 2] += gy * c4; f[fw * i + 3] += gz * c4;
    }
    if (field & 4)
      for (j = 0; j < 6; j++) f[fw * i + 4 + j] += h[j] * c4;
  }
}

/* ======================================================================
 * octree
 * ====================================================================== */

typedef struct {
  int depth, nsrc;
  double x0[3], D;
  uint32_t *ip;          /* Morton-sorted source indices */
  uint64_t *code;        /* Morton codes of sorted sources */
  int *bstart[17], *bcount[17];  /* dense per level */
} orc_tree;

typedef struct { uint64_t c; uint32_t i; } code_idx;

static int cmp_code(const void *a, const void *b)
{
  const code_idx *x = a, *y = b;
  if (x->c < y->c) return -1;
  if (x->c > y->c) return 1;
  /* stable: ties keep input order (GLib qsort is stable, src/boxes.c:158) */
  return (x->i < y->i) ? -1 : (x->i > y->i);
}

/* tree build: src/boxes.c:118-277 (sort by Morton code, box ranges) */
static int tree_build(orc_tree *t, int nsrc, const double *xs)
{
  int i, l;
  code_idx *ci = malloc(sizeof(code_idx) * (nsrc > 0 ? nsrc : 1));
  for (i = 0; i < nsrc; i++) {
    const double *x = xs + 3 * i;
    if (x[0] <= t->x0[0] || x[0] >= t->x0[0] + t->D ||
        x[1] <= t->x0[1] || x[1] >= t->x0[1] + t->D ||
        x[2] <= t->x0[2] || x[2] >= t->x0[2] + t->D) {
      free(ci);
      return -1;
    }
    ci[i].c = orc_point_index(x, t->x0, t->D);
    ci[i].i = i;
  }
  qsort(ci, nsrc, sizeof(code_idx), cmp_code);
  t->nsrc = nsrc;
  t->ip = malloc(sizeof(uint32_t) * (nsrc + 1));
  t->code = malloc(sizeof(uint64_t) * (nsrc + 1));
  for (i = 0; i < nsrc; i++) { t->ip[i] = ci[i].i; t->code[i] = ci[i].c; }
  free(ci);
  for (l = 0; l <= t->depth; l++) {
    int nb = 1 << (3 * l);
    t->bstart[l] = calloc(nb, sizeof(int));
    t->bcount[l] = calloc(nb, sizeof(int));
    for (i = 0; i < nsrc; i++) {
      uint64_t b = t->code[i] >> (3 * (20 - l));
      if (t->bcount[l][b] == 0) t->bstart[l][b] = i;
      t->bcount[l][b]++;
    }
  }
  return 0;
}

static void tree_free(orc_tree *t)
{
  int l;
  free(t->ip); free(t->code);
  for (l = 0; l <= t->depth; l++) { free(t->bstart[l]); free(t->bcount[l]); }
}

static void box_centre(const orc_tree *t, int level, uint64_t b, double *xc)
{
  uint32_t i, j, k;
  double w = t->D / (1 << level);
  orc_morton_decode(b, &i, &j, &k);
  xc[0] = t->x0[0] + w * (i + 0.5);
  xc[1] = t->x0[1] + w * (j + 0.5);
  xc[2] = t->x0[2] + w * (k + 0.5);
}

/* ======================================================================
 * Laplace solve: src/laplace.c:854-989 (P2M), src/pass-laplace.c:433-486
 * with src/shift-laplace.c:225-442 (M2M), src/pass-laplace.c:35-408 (M2L)
 * with src/shift-laplace.c:562-789 (L2L), src/local-laplace.c:1293-1336
 * (L2P + near field); call order of tools/wbfmm-laplace-tree-calc.c.in
 * :380-460.
 * ====================================================================== */


/* ======================================================================
 * Helmholtz operators.  Coefficients complex, C_n^m at 2*HIDX(n,m),
 * -n<=m<=n (src/wbfmm.h:242).  Kernel G = h0(kr)/(4pi) = e^{ikr}/(4pi ikr)
 * (src/direct.c:58-70).
 * ====================================================================== */

#include <complex.h>
typedef double complex cplx;

/* R_n^m = j_n(kr) Pbar_n^|m| e^{i m phi} for n<=N, all m, plus (when grad)
   the ingredients for n<=N+1 */
static void hel_regular(double k, int N, const double *x0, const double *x,
                        cplx *R)
{
  double r, C, S, cp, sp, P[(66 + 1) * (66 + 2) / 2], j[70];
  int n, m;
  sph(x0, x, &r, &C, &S, &cp, &sp);
  orc_legendre(N, C, S, P);
  orc_bessel_j(N, k * r, j);
  for (n = 0; n <= N; n++)
    for (m = 0; m <= n; m++) {
      cplx e = cpow(cp + I * sp, m);
      double a = j[n] * P[PIDX(n, m)];
      R[HIDX(n, m)] = a * e;
      R[HIDX(n, -m)] = a * conj(e);
    }
}

/* P2M monopole (src/expansions.c:52-123): M_n^m += q conj(R_n^m(xs-x0)) */
void orc_hel_p2m(double k, int N, const double *x0, const double *xs,
                 const double *q, double *M)
{
  cplx R[(66 + 1) * (66 + 1)], qq = q[0] + I * q[1];
  int i;
  hel_regular(k, N, x0, xs, R);
  for (i = 0; i < (N + 1) * (N + 1); i++) {
    cplx v = qq * conj(R[i]);
    M[2 * i] += creal(v);
    M[2 * i + 1] += cimag(v);
  }
}

/* derivative coefficients of the regular wave functions, G&D (2004)
   eq. 3.7 (used by src/gradient.c:220-280):
     dz R_n^m = k [a_{n-1}^m R_{n-1}^m - a_n^m R_{n+1}^m]
     D+ R_n^m = k [b_{n+1}^{-m-1} R_{n+1}^{m+1} - b_n^m R_{n-1}^{m+1}]
     D- R_n^m = k [b_{n+1}^{m-1} R_{n+1}^{m-1} - b_n^{-m} R_{n-1}^{m-1}]
   with a, b the recursion coefficients of src/translations.c:182-203 */
static double dab_a(int n, int m) { return (double)rec_a(n, m); }
static double dab_b(int n, int m) { return (double)rec_b(n, m); }

static cplx Rget(const cplx *R, int N, int n, int m)
{
  if (n < 0 || n > N || m < -n || m > n) return 0.0;
  return R[HIDX(n, m)];
}

/* gradient of R_n^m at the point whose R table (to order N+1) is given */
static void hel_regular_grad(double k, int N1, const cplx *R, int n, int m,
                             cplx *gx, cplx *gy, cplx *gz)
{
  cplx dz = k * (dab_a(n - 1, m) * Rget(R, N1, n - 1, m) - dab_a(n, m) * Rget(R, N1, n + 1, m));
  cplx dp = k * (dab_b(n + 1, -m - 1) * Rget(R, N1, n + 1, m + 1) - dab_b(n, m) * Rget(R, N1, n - 1, m + 1));
  cplx dm = k * (dab_b(n + 1, m - 1) * Rget(R, N1, n + 1, m - 1) - dab_b(n, -m) * Rget(R, N1, n - 1, m - 1));
  *gz = dz;
  *gx = 0.5 * (dp + dm);
  *gy = -0.5 * I * (dp - dm);
}

/* P2M normal dipole: the source d nrm.grad_x G = -d nrm.grad_y G
   (src/local.c:155-166) expands as M_n^m += -d conj(nrm . grad R_n^m(xs-x0))
   (src/expansions.c:545-631) */
void orc_hel_p2m_normal(double k, int N, const double *x0, const double *xs,
                        const double *nrm, const double *d, double *M)
{
  cplx R[(67 + 1) * (67 + 1)], dd = d[0] + I * d[1];
  int n, m;
  hel_regular(k, N + 1, x0, xs, R);
  for (n = 0; n <= N; n++)
    for (m = -n; m <= n; m++) {
      cplx gx, gy, gz, v;
      hel_regular_grad(k, N + 1, R, n, m, &gx, &gy, &gz);
      v = -dd * conj(nrm[0] * gx + nrm[1] * gy + nrm[2] * gz);
      M[2 * HIDX(n, m)] += creal(v);
      M[2 * HIDX(n, m) + 1] += cimag(v);
    }
}

/* singular evaluation (src/expansions.c:147-228):
   f += sum M_n^m h_n(kr) Pbar_n^|m| e^{i m phi} */
void orc_hel_h_eval(double k, int N, const double *x0, const double *M,
                    const double *xf, double *f)
{
  double r, C, S, cp, sp, P[(66 + 1) * (66 + 2) / 2], h[2 * 70];
  int n, m;
  cplx acc = 0.0;
  sph(x0, xf, &r, &C, &S, &cp, &sp);
  orc_legendre(N, C, S, P);
  orc_bessel_h(N, k * r, h);
  for (n = 0; n <= N; n++) {
    cplx hn = h[2 * n] + I * h[2 * n + 1];
    for (m = -n; m <= n; m++) {
      cplx e = cpow(cp + I * sp, m);
      acc += (M[2 * HIDX(n, m)] + I * M[2 * HIDX(n, m) + 1]) * hn * P[PIDX(n, abs(m))] * e;
    }
  }
  f[0] += creal(acc);
  f[1] += cimag(acc);
}

/* regular evaluation (L2P): field&1 potential f[0..1], field&2 gradient
   f[2..7] (src/expansions.c:275-344, src/gradient.c:282-399) */
void orc_hel_j_eval(double k, int N, const double *x0, const double *L,
                    const double *xf, int field, double *f)
{
  cplx R[(67 + 1) * (67 + 1)];
  cplx pot = 0, gx = 0, gy = 0, gz = 0;
  int n, m;
  hel_regular(k, N + 1, x0, xf, R);
  for (n = 0; n <= N; n++)
    for (m = -n; m <= n; m++) {
      cplx c = L[2 * HIDX(n, m)] + I * L[2 * HIDX(n, m) + 1];
      if (field & 1) pot += c * R[HIDX(n, m)];
      if (field & 2) {
        cplx ax, ay, az;
        hel_regular_grad(k, N + 1, R, n, m, &ax, &ay, &az);
        gx += c * ax; gy += c * ay; gz += c * az;
      }
    }
  if (field & 1) { f[0] += creal(pot); f[1] += cimag(pot); }
  if (field & 2) {
    f[2] += creal(gx); f[3] += cimag(gx);
    f[4] += creal(gy); f[5] += cimag(gy);
    f[6] += creal(gz); f[7] += cimag(gz);
  }
}

#define CIDX(l, m, n) ((l) * ((l) + 1) * ((l) + 2) / 6 + (n) * ((n) + 1) / 2 + (m))

/*
 * coaxial coefficients (E|F)_{l n}^m, l >= n >= m >= 0, at distance kr
 * (kind 0: R|R = S|S from j_l, kind 1: S|R from h_l), G&D 4.79/4.83/4.84,
 * restating src/translations.c:205-409; computed in quad precision (the
 * recursion runs to degree 2L).  E holds 2 doubles (re, im) per entry,
 * entries CIDX(l,m,n) for l <= L.
 */
void orc_hel_coaxial_table(int kind, int L, double kr, double *E)
{
  int l, m, n, sw = 1, LL = 2 * L;
  size_t ne = (size_t)CIDX(LL + 1, 0, 0) + 8;
  qreal *er = calloc(ne, sizeof(qreal)), *ei = calloc(ne, sizeof(qreal));
  qreal x, f0r[2 * 64 + 4], f0i[2 * 64 + 4];
  if (kr < 0) { kr = -kr; sw = -1; }
  x = (qreal)kr;
  /* f_l(x): j_l (kind 0) or h_l = j_l + i y_l (kind 1), quad precision,
     j_l by downward recursion normalised to j_0, y_l upward */
  {
    int top = LL + 40;
    qreal jp = 0, jc = 1e-300Q, jm, *jj = calloc(top + 2, sizeof(qreal));
    qreal y0 = -cosq(x) / x, y1 = -cosq(x) / (x * x) - sinq(x) / x, ym, yc, yn;
    int i;
    for (i = top; i >= 1; i--) {
      jm = (2 * i + 1) / x * jc - jp;
      jp = jc; jc = jm;
      jj[i - 1] = jc;
      if (fabsq(jc) > 1e250Q) {
        int t2;
        for (t2 = i - 1; t2 <= top; t2++) jj[t2] *= 1e-250Q;
        jp *= 1e-250Q; jc *= 1e-250Q;
      }
    }
    {
      qreal scale = (sinq(x) / x) / jj[0];
      for (i = 0; i <= LL + 1; i++) f0r[i] = jj[i] * scale;
    }
    free(jj);
    ym = y0; yc = y1;
    f0i[0] = y0;
    if (LL >= 1) f0i[1] = y1;
    for (i = 1; i < LL + 1; i++) {
      yn = (2 * i + 1) / x * yc - ym;
      ym = yc; yc = yn;
      f0i[i + 1] = yc;
    }
  }
  for (l = 0; l <= LL; l++) {
    qreal sg = (sw < 0 && (l % 2)) ? -1 : 1;
    er[CIDX(l, 0, 0)] = sg * sqrtq((qreal)(2 * l + 1)) * f0r[l];
    ei[CIDX(l, 0, 0)] = kind == 1 ? sg * sqrtq((qreal)(2 * l + 1)) * f0i[l] : 0;
  }
  /* G&D 4.79, m = 0: (E|F)_{l,n+1}^0 from n-1, n and l+-1 */
  m = 0;
  for (n = 0; n <= L; n++)
    for (l = n + 1; l <= LL - n; l++) {
      qreal a1 = rec_a(n - 1, m), a2 = rec_a(n, m), a3 = rec_a(l, m), a4 = rec_a(l - 1, m);
      qreal e1r = n >= 1 ? er[CIDX(l, m, n - 1)] : 0, e1i = n >= 1 ? ei[CIDX(l, m, n - 1)] : 0;
      er[CIDX(l, m, n + 1)] = (a1 * e1r - a3 * er[CIDX(l + 1, m, n)] + a4 * er[CIDX(l - 1, m, n)]) / a2;
      ei[CIDX(l, m, n + 1)] = (a1 * e1i - a3 * ei[CIDX(l + 1, m, n)] + a4 * ei[CIDX(l - 1, m, n)]) / a2;
    }
  /* G&D 4.84: (E|F)_{l,m+1}^{m+1} */
  for (m = 0; m <= L; m++)
    for (l = m + 1; l <= LL - m; l++) {
      qreal a1 = rec_b(m + 1, -m - 1), a2 = rec_b(l, -m - 1), a3 = rec_b(l + 1, m);
      er[CIDX(l, m + 1, m + 1)] = (a2 * er[CIDX(l - 1, m, m)] - a3 * er[CIDX(l + 1, m, m)]) / a1;
      ei[CIDX(l, m + 1, m + 1)] = (a2 * ei[CIDX(l - 1, m, m)] - a3 * ei[CIDX(l + 1, m, m)]) / a1;
    }
  /* G&D 4.79 for m >= 1 */
  for (m = 1; m <= L; m++)
    for (n = m; n <= L; n++)
      for (l = n + 1; l <= LL - n; l++) {
        qreal a1 = rec_a(n - 1, m), a2 = rec_a(n, m), a3 = rec_a(l, m), a4 = rec_a(l - 1, m);
        qreal e1r = (n - 1 >= m) ? er[CIDX(l, m, n - 1)] : 0;
        qreal e1i = (n - 1 >= m) ? ei[CIDX(l, m, n - 1)] : 0;
        er[CIDX(l, m, n + 1)] = (a1 * e1r - a3 * er[CIDX(l + 1, m, n)] + a4 * er[CIDX(l - 1, m, n)]) / a2;
        ei[CIDX(l, m, n + 1)] = (a1 * e1i - a3 * ei[CIDX(l + 1, m, n)] + a4 * ei[CIDX(l - 1, m, n)]) / a2;
      }
  for (l = 0; l <= L; l++)
    for (n = 0; n <= l; n++)
      for (m = 0; m <= n; m++) {
        E[2 * CIDX(l, m, n)] = (double)er[CIDX(l, m, n)];
        E[2 * CIDX(l, m, n) + 1] = (double)ei[CIDX(l, m, n)];
      }
  free(er);
  free(ei);
}

/*
 * coaxial translation (src/translations.c:411-604):
 *   Co_n^m = sc Co + sum_l T(n,l,|m|) Ci_l^m,
 *   T = E_{l n}^{|m|} (l >= n), (-1)^{n+l} E_{n l}^{|m|} (l < n)
 */
void orc_hel_coaxial(double *Co, int No, const double *Ci, int Ni,
                     const double *E, int L, int is_complex, double sc)
{
  int n, m, l;
  (void)L;
  for (n = 0; n <= No; n++)
    for (m = -n; m <= n; m++) {
      int am = abs(m);
      cplx acc = 0;
      for (l = am; l <= Ni; l++) {
        int idx = (l >= n) ? CIDX(l, am, n) : CIDX(n, am, l);
        double sg = (l < n && ((n + l) % 2)) ? -1.0 : 1.0;
        cplx e = E[2 * idx] + I * (is_complex ? E[2 * idx + 1] : 0.0);
        acc += sg * e * (Ci[2 * HIDX(l, m)] + I * Ci[2 * HIDX(l, m) + 1]);
      }
      if (sc == 0.0) {
        Co[2 * HIDX(n, m)] = creal(acc);
        Co[2 * HIDX(n, m) + 1] = cimag(acc);
      } else {
        Co[2 * HIDX(n, m)] = sc * Co[2 * HIDX(n, m)] + creal(acc);
        Co[2 * HIDX(n, m) + 1] = sc * Co[2 * HIDX(n, m) + 1] + cimag(acc);
      }
    }
}

/* rotation (src/rotations.c:419-593):
   Co_n^nu = sc Co + e^{-i nu ph} sum_m H_n^{nu,m} e^{i m ch} Ci_n^m,
   H_n^{nu,-m} = H_n^{-nu,m} */
void orc_hel_rotate(double *Co, const double *Ci, int N, const double *H,
                    double ph, double ch, double sc)
{
  int n, m, nu;
  for (n = 0; n <= N; n++)
    for (nu = -n; nu <= n; nu++) {
      cplx acc = 0;
      for (m = -n; m <= n; m++) {
        double h = (m >= 0) ? H[RIDX(nu, m, n)] : H[RIDX(-nu, -m, n)];
        acc += h * cexp(I * m * ch) * (Ci[2 * HIDX(n, m)] + I * Ci[2 * HIDX(n, m) + 1]);
      }
      acc *= cexp(-I * nu * ph);
      if (sc == 0.0) {
        Co[2 * HIDX(n, nu)] = creal(acc);
        Co[2 * HIDX(n, nu) + 1] = cimag(acc);
      } else {
        Co[2 * HIDX(n, nu)] = sc * Co[2 * HIDX(n, nu)] + creal(acc);
        Co[2 * HIDX(n, nu) + 1] = sc * Co[2 * HIDX(n, nu) + 1] + cimag(acc);
      }
    }
}

/* FP64 direct sum (src/direct.c:35-294), G = h0(kr)/(4 pi), complex
   monopoles q and normal dipoles d: f[8i..8i+7] = potential, gradient
   (re,im pairs).  Pairs with r <= cutoff skipped. */
void orc_hel_direct(double k, int nsrc, const double *xs, const double *q,
                    const double *nrm, const double *d,
                    int ntgt, const double *xt, int field, double cutoff,
                    double *f)
{
  int i, j;
  for (i = 0; i < ntgt; i++) {
    cplx pot = 0, g[3] = {0, 0, 0};
    for (j = 0; j < nsrc; j++) {
      double rv[3] = {xt[3 * i] - xs[3 * j], xt[3 * i + 1] - xs[3 * j + 1],
                      xt[3 * i + 2] - xs[3 * j + 2]};
      double r = sqrt(rv[0] * rv[0] + rv[1] * rv[1] + rv[2] * rv[2]);
      cplx G, dG, ddG;
      int a;
      if (r <= cutoff) continue;
      /* G(r) = e^{ikr}/(4 pi i k r); G' = dG/dr; G'' */
      G = cexp(I * k * r) / (4.0 * M_PI * I * k * r);
      dG = G * (I * k - 1.0 / r);
      ddG = G * ((I * k - 1.0 / r) * (I * k - 1.0 / r) + 1.0 / (r * r));
      if (q) {
        cplx qq = q[2 * j] + I * q[2 * j + 1];
        pot += qq * G;
        for (a = 0; a < 3; a++) g[a] += qq * dG * rv[a] / r;
      }
      if (d) {
        cplx dd = d[2 * j] + I * d[2 * j + 1];
        const double *nv = nrm + 3 * j;
        double nr = (nv[0] * rv[0] + nv[1] * rv[1] + nv[2] * rv[2]) / r;
        /* the Helmholtz dipole is d n.grad_x G = d G'(r) n.(x-y)/r, i.e.
           -k h1(kr) d n.(x-y)/(4 pi r) (src/local.c:155-166) */
        pot += dd * dG * nr;
        for (a = 0; a < 3; a++) {
          double ra = rv[a] / r;
          g[a] += dd * (ddG * nr * ra + dG * (nv[a] - nr * ra) / r);
        }
      }
    }
    if (field & 1) { f[8 * i] += creal(pot); f[8 * i + 1] += cimag(pot); }
    if (field & 2) {
      int a;
      for (a = 0; a < 3; a++) { f[8 * i + 2 + 2 * a] += creal(g[a]); f[8 * i + 3 + 2 * a] += cimag(g[a]); }
    }
  }
}

#include <pthread.h>

typedef struct {
  const orc_tree *t;
  const orc_fmm_cfg *cfg;
  int level, nth, ith;
  double **mps, **mpr;
  double *Htab;   /* per (dx,dy,dz) offset rotation tables */
  int nrot;
} m2l_job;

static void *lap_m2l_worker(void *arg)
{
  m2l_job *J = arg;
  const orc_tree *t = J->t;
  int l = J->level, Ns = J->cfg->order_s[l], Nr = J->cfg->order_r[l];
  int ncs = (Ns + 1) * (Ns + 1), ncr = (Nr + 1) * (Nr + 1);
  uint64_t nb = 1ull << (3 * l), b;
  double w = t->D / (1 << l);
  double *wks = malloc(sizeof(double) * ncs), *wkr = malloc(sizeof(double) * ncr);
  uint64_t list[216];
  int off[3 * 216];
  for (b = J->ith; b < nb; b += J->nth) {
    int n = orc_box_list4(l, b, list, off), j;
    double *L = J->mpr[l] + b * ncr;
    for (j = 0; j < n; j++) {
      double th, ph, ch, r;
      const double *H;
      if (t->bcount[l][list[j]] == 0) continue;  /* zero expansion */
      orc_shift_angles(off[3 * j], off[3 * j + 1], off[3 * j + 2], &th, &ph, &ch, &r);
      H = J->Htab + (size_t)((off[3 * j] + 3) * 49 + (off[3 * j + 1] + 3) * 7 +
                             off[3 * j + 2] + 3) * J->nrot;
      orc_lap_rotate(wks, J->mps[l] + list[j] * ncs, Ns, H, ph, ch, 0.0);
      orc_lap_coaxial(2, wkr, Nr, wks, Ns, r * w, 0.0);
      orc_lap_rotate(L, wkr, Nr, H, ch, ph, 1.0);
    }
  }
  free(wks); free(wkr);
  return NULL;
}

static void run_m2l(orc_tree *t, orc_fmm_cfg *cfg, int level, double **mps,
                    double **mpr, double *Htab, int nrot,
                    void *(*worker)(void *))
{
  int nth = cfg->nthreads > 1 ? cfg->nthreads : 1, i;
  pthread_t th[64];
  m2l_job jobs[64];
  if (nth > 64) nth = 64;
  for (i = 0; i < nth; i++) {
    jobs[i].t = t; jobs[i].cfg = cfg; jobs[i].level = level;
    jobs[i].nth = nth; jobs[i].ith = i; jobs[i].mps = mps; jobs[i].mpr = mpr;
    jobs[i].Htab = Htab; jobs[i].nrot = nrot;
  }
  if (nth == 1) { worker(&jobs[0]); return; }
  for (i = 0; i < nth; i++) pthread_create(&th[i], NULL, worker, &jobs[i]);
  for (i = 0; i < nth; i++) pthread_join(th[i], NULL);
}

/* rotation tables for all 343 offsets up to order N (entries for
   offsets in the near field are left unused) */
static double *offset_rotations(int N, int *nrot)
{
  int dx, dy, dz;
  double *Htab;
  *nrot = orc_rotation_size(N);
  Htab = calloc((size_t)343 * (*nrot), sizeof(double));
  for (dx = -3; dx <= 3; dx++)
    for (dy = -3; dy <= 3; dy++)
      for (dz = -3; dz <= 3; dz++) {
        double th, ph, ch, r;
        if (abs(dx) <= 1 && abs(dy) <= 1 && abs(dz) <= 1) continue;
        orc_shift_angles(dx, dy, dz, &th, &ph, &ch, &r);
        orc_H_rotation(Htab + (size_t)((dx + 3) * 49 + (dy + 3) * 7 + dz + 3) * (*nrot),
                       N, th);
      }
  return Htab;
}

/* child/parent direction angles: child Morton digit c, direction from
   child centre to parent centre (sgn=+1) or reverse (sgn=-1) */
static void child_angles(int c, int sgn, double *th, double *ph, double *ch)
{
  double ex = sgn * (0.5 - (c & 1)), ey = sgn * (0.5 - ((c >> 1) & 1));
  double ez = sgn * (0.5 - ((c >> 2) & 1));
  double r = sqrt(ex * ex + ey * ey + ez * ez);
  *th = acos(ez / r);
  *ph = -0.5 * M_PI;
  *ch = atan2(ey, ex);
}

static int lap_solve(orc_fmm_cfg *cfg, orc_tree *t, const double *xs,
                     const double *q, const double *nrm, const double *d,
                     int ntgt, const double *xt, double *f)
{
  int depth = t->depth, l, Lmax = 0, nrot, i;
  double *mps[17] = {0}, *mpr[17] = {0}, *Htab, tm;
  double *Hud[2][8];
  if ((cfg->field & 4) && d) return -4;  /* second derivatives: monopoles only */
  for (l = 2; l <= depth; l++) {
    if (cfg->order_s[l] > Lmax) Lmax = cfg->order_s[l];
    if (cfg->order_r[l] > Lmax) Lmax = cfg->order_r[l];
  }
  Htab = offset_rotations(Lmax, &nrot);
  for (i = 0; i < 8; i++) {
    double th, ph, ch;
    Hud[0][i] = malloc(sizeof(double) * nrot);
    Hud[1][i] = malloc(sizeof(double) * nrot);
    child_angles(i, 1, &th, &ph, &ch); orc_H_rotation(Hud[0][i], Lmax, th);
    child_angles(i, -1, &th, &ph, &ch); orc_H_rotation(Hud[1][i], Lmax, th);
  }
  for (l = 2; l <= depth; l++) {
    size_t nb = (size_t)1 << (3 * l);
    mps[l] = calloc(nb * (cfg->order_s[l] + 1) * (cfg->order_s[l] + 1), sizeof(double));
    mpr[l] = calloc(nb * (cfg->order_r[l] + 1) * (cfg->order_r[l] + 1), sizeof(double));
  }
  /* P2M at the leaves */
  tm = now_s();
  {
    int Ns = cfg->order_s[depth], ncs = (Ns + 1) * (Ns + 1);
    uint64_t nb = 1ull << (3 * depth), b;
    for (b = 0; b < nb; b++) {
      double xc[3];
      int j;
      if (t->bcount[depth][b] == 0) continue;
      box_centre(t, depth, b, xc);
      for (j = 0; j < t->bcount[depth][b]; j++) {
        int s = t->ip[t->bstart[depth][b] + j];
        if (d) orc_lap_p2m_normal(Ns, xc, xs + 3 * s, nrm + 3 * s, d[s], mps[depth] + b * ncs);
        if (q) orc_lap_p2m(Ns, xc, xs + 3 * s, q[s], mps[depth] + b * ncs);
      }
    }
  }
  cfg->t_phase[1] = now_s() - tm;
  /* upward pass, levels depth..3 */
  tm = now_s();
  for (l = depth; l >= 3; l--) {
    int Nc = cfg->order_s[l], Np = cfg->order_s[l - 1];
    int ncc = (Nc + 1) * (Nc + 1), ncp = (Np + 1) * (Np + 1);
    double wc = t->D / (1 << l), tt = sqrt(3.0) * 0.5 * wc;
    uint64_t np = 1ull << (3 * (l - 1)), p;
    double *wk1 = malloc(sizeof(double) * (ncc > ncp ? ncc : ncp));
    double *wk2 = malloc(sizeof(double) * ncp);
    for (p = 0; p < np; p++) {
      int c;
      double *Cp = mps[l - 1] + p * ncp;
      memset(Cp, 0, sizeof(double) * ncp);
      if (t->bcount[l - 1][p] == 0) continue;
      for (c = 0; c < 8; c++) {
        double th, ph, ch;
        if (t->bcount[l][8 * p + c] == 0) continue;
        child_angles(c, 1, &th, &ph, &ch);
        orc_lap_rotate(wk1, mps[l] + (8 * p + c) * ncc, Nc, Hud[0][c], ph, ch, 0.0);
        orc_lap_coaxial(0, wk2, Np, wk1, Nc, tt, 0.0);
        orc_lap_rotate(Cp, wk2, Np, Hud[0][c], ch, ph, 1.0);
      }
    }
    free(wk1); free(wk2);
  }
  cfg->t_phase[2] = now_s() - tm;
  /* downward pass, levels 2..depth */
  tm = now_s();
  for (l = 2; l <= depth; l++) {
    run_m2l(t, cfg, l, mps, mpr, Htab, nrot, lap_m2l_worker);
    if (l < depth) {
      int Np = cfg->order_r[l], Nc = cfg->order_r[l + 1];
      int ncp = (Np + 1) * (Np + 1), ncc = (Nc + 1) * (Nc + 1);
      double wc = t->D / (1 << (l + 1)), tt = sqrt(3.0) * 0.5 * wc;
      uint64_t np = 1ull << (3 * l), p;
      double *wk1 = malloc(sizeof(double) * ncp), *wk2 = malloc(sizeof(double) * ncc);
      for (p = 0; p < np; p++) {
        int c;
        for (c = 0; c < 8; c++) {
          double th, ph, ch;
          child_angles(c, -1, &th, &ph, &ch);
          orc_lap_rotate(wk1, mpr[l] + p * ncp, Np, Hud[1][c], ph, ch, 0.0);
          orc_lap_coaxial(1, wk2, Nc, wk1, Np, tt, 0.0);
          orc_lap_rotate(mpr[l + 1] + (8 * p + c) * ncc, wk2, Nc, Hud[1][c], ch, ph, 1.0);
        }
      }
      free(wk1); free(wk2);
    }
  }
  cfg->t_phase[3] = now_s() - tm;
  /* L2P + near field */
  tm = now_s();
  {
    int Nr = cfg->order_r[depth], ncr = (Nr + 1) * (Nr + 1);
    for (i = 0; i < ntgt; i++) {
      const double *x = xt + 3 * i;
      uint64_t b = orc_point_index(x, t->x0, t->D) >> (3 * (20 - depth));
      uint64_t nb[27];
      double xc[3], loc[4] = {0, 0, 0, 0}, hes[6] = {0, 0, 0, 0, 0, 0};
      int nn, a, j;
      box_centre(t, depth, b, xc);
      orc_lap_l_eval(Nr, xc, mpr[depth] + b * ncr, x, cfg->field, loc);
      if (cfg->field & 4) orc_lap_l_hess(Nr, xc, mpr[depth] + b * ncr, x, hes);
      nn = orc_box_neighbours(depth, b, nb);
      for (a = 0; a < nn; a++) {
        for (j = 0; j < t->bcount[depth][nb[a]]; j++) {
          int s = t->ip[t->bstart[depth][nb[a]] + j];
          double rx = x[0] - xs[3 * s], ry = x[1] - xs[3 * s + 1], rz = x[2] - xs[3 * s + 2];
          double r2 = rx * rx + ry * ry + rz * rz, r, g;
          if (r2 <= CUTOFF * CUTOFF) continue;
          r = sqrt(r2);
          g = 0.25 / M_PI / r;
          if (q) {
            loc[0] += q[s] * g;
            loc[1] -= q[s] * g * rx / r2; loc[2] -= q[s] * g * ry / r2;
            loc[3] -= q[s] * g * rz / r2;
            if (cfg->field & 4) {
              /* d_j d_k G = (3 r_j r_k / r^2 - delta_jk) G / r^2 */
              const double h = q[s] * g / r2;
              hes[0] += h * (3.0 * rx * rx / r2 - 1.0);
              hes[1] += h * (3.0 * ry * ry / r2 - 1.0);
              hes[2] += h * (3.0 * rz * rz / r2 - 1.0);
              hes[3] += h * 3.0 * rx * ry / r2;
              hes[4] += h * 3.0 * ry * rz / r2;
              hes[5] += h * 3.0 * rz * rx / r2;
            }
          }
          if (d) {
            const double *nv = nrm + 3 * s;
            double nr = nv[0] * rx + nv[1] * ry + nv[2] * rz;
            loc[0] += d[s] * nr * g / r2;
            loc[1] += d[s] * g / r2 * (nv[0] - 3.0 * nr * rx / r2);
            loc[2] += d[s] * g / r2 * (nv[1] - 3.0 * nr * ry / r2);
            loc[3] += d[s] * g / r2 * (nv[2] - 3.0 * nr * rz / r2);
          }
        }
      }
      {
        /* rows of 4 (potential, gradient) or, with field bit 4, of 10
           (and the second derivatives xx, yy, zz, xy, yz, zx) */
        const int fw = (cfg->field & 4) ? 10 : 4;
        if (cfg->field & 1) f[fw * i] += loc[0];
        if (cfg->field & 2) { f[fw * i + 1] += loc[1]; f[fw * i + 2] += loc[2]; f[fw * i + 3] += loc[3]; }
        if (cfg->field & 4)
          for (a = 0; a < 6; a++) f[fw * i + 4 + a] += hes[a];
      }
    }
  }
  cfg->t_phase[4] = now_s() - tm;
  if (cfg->dump_level >= 2 && cfg->dump_level <= depth) {
    int l2 = cfg->dump_level;
    size_t nb = (size_t)1 << (3 * l2);
    if (cfg->dump_mps)
      memcpy(cfg->dump_mps, mps[l2], sizeof(double) * nb * (cfg->order_s[l2] + 1) * (cfg->order_s[l2] + 1));
    if (cfg->dump_mpr)
      memcpy(cfg->dump_mpr, mpr[l2], sizeof(double) * nb * (cfg->order_r[l2] + 1) * (cfg->order_r[l2] + 1));
  }
  for (l = 2; l <= depth; l++) { free(mps[l]); free(mpr[l]); }
  for (i = 0; i < 8; i++) { free(Hud[0][i]); free(Hud[1][i]); }
  free(Htab);
  return 0;
}


/* ======================================================================
 * Helmholtz solve: src/boxes.c:341-490 (P2M), src/pass.c:581-632 with
 * src/shift.c:244-401 (M2M), src/pass.c:56-105, :448-537 (M2L),
 * src/shift.c:403-585 (L2L), src/local.c:32-262 (L2P + near field); call
 * order of tools/wbfmm-tree-calc.c.in:306-417.  Translation tables per
 * level: S|R at the 15 list-4 distances k w r_i, R|R (= S|S) at
 * k w sqrt(3)/2 (src/shift.c:777-861).
 * ====================================================================== */

typedef struct {
  const orc_tree *t;
  const orc_fmm_cfg *cfg;
  int level, nth, ith;
  double **mps, **mpr;
  double *Htab;
  int nrot;
  double **SR;      /* per offset slot (343), table or NULL */
} hel_job;

static void *hel_m2l_worker(void *arg)
{
  hel_job *J = arg;
  const orc_tree *t = J->t;
  int l = J->level, Ns = J->cfg->order_s[l], Nr = J->cfg->order_r[l];
  int L = Ns > Nr ? Ns : Nr;
  int ncs = 2 * (Ns + 1) * (Ns + 1), ncr = 2 * (Nr + 1) * (Nr + 1);
  uint64_t nb = 1ull << (3 * l), b;
  double *wks = malloc(sizeof(double) * 2 * (L + 1) * (L + 1));
  double *wkr = malloc(sizeof(double) * 2 * (L + 1) * (L + 1));
  uint64_t list[216];
  int off[3 * 216];
  for (b = J->ith; b < nb; b += J->nth) {
    int n = orc_box_list4(l, b, list, off), j;
    double *Lc = J->mpr[l] + b * ncr;
    for (j = 0; j < n; j++) {
      double th, ph, ch, r;
      int o = (off[3 * j] + 3) * 49 + (off[3 * j + 1] + 3) * 7 + off[3 * j + 2] + 3;
      if (t->bcount[l][list[j]] == 0) continue;
      orc_shift_angles(off[3 * j], off[3 * j + 1], off[3 * j + 2], &th, &ph, &ch, &r);
      orc_hel_rotate(wks, J->mps[l] + list[j] * ncs, Ns, J->Htab + (size_t)o * J->nrot, ph, ch, 0.0);
      orc_hel_coaxial(wkr, Nr, wks, Ns, J->SR[o], L, 1, 0.0);
      orc_hel_rotate(Lc, wkr, Nr, J->Htab + (size_t)o * J->nrot, ch, ph, 1.0);
    }
  }
  free(wks); free(wkr);
  return NULL;
}

static int hel_solve(orc_fmm_cfg *cfg, orc_tree *t, const double *xs,
                     const double *q, const double *nrm, const double *d,
                     int ntgt, const double *xt, double *f)
{
  int depth = t->depth, l, Lmax = 0, nrot, i;
  double *mps[17] = {0}, *mpr[17] = {0}, *Htab, tm, k = cfg->k;
  double *Hud[2][8];
  for (l = 2; l <= depth; l++) {
    if (cfg->order_s[l] > Lmax) Lmax = cfg->order_s[l];
    if (cfg->order_r[l] > Lmax) Lmax = cfg->order_r[l];
  }
  Htab = offset_rotations(Lmax, &nrot);
  for (i = 0; i < 8; i++) {
    double th, ph, ch;
    Hud[0][i] = malloc(sizeof(double) * nrot);
    Hud[1][i] = malloc(sizeof(double) * nrot);
    child_angles(i, 1, &th, &ph, &ch); orc_H_rotation(Hud[0][i], Lmax, th);
    child_angles(i, -1, &th, &ph, &ch); orc_H_rotation(Hud[1][i], Lmax, th);
  }
  for (l = 2; l <= depth; l++) {
    size_t nb = (size_t)1 << (3 * l);
    mps[l] = calloc(nb * 2 * (cfg->order_s[l] + 1) * (cfg->order_s[l] + 1), sizeof(double));
    mpr[l] = calloc(nb * 2 * (cfg->order_r[l] + 1) * (cfg->order_r[l] + 1), sizeof(double));
  }
  tm = now_s();
  {
    int Ns = cfg->order_s[depth], ncs = 2 * (Ns + 1) * (Ns + 1);
    uint64_t nb = 1ull << (3 * depth), b;
    for (b = 0; b < nb; b++) {
      double xc[3];
      int j;
      if (t->bcount[depth][b] == 0) continue;
      box_centre(t, depth, b, xc);
      for (j = 0; j < t->bcount[depth][b]; j++) {
        int s = t->ip[t->bstart[depth][b] + j];
        if (q) orc_hel_p2m(k, Ns, xc, xs + 3 * s, q + 2 * s, mps[depth] + b * ncs);
        if (d) orc_hel_p2m_normal(k, Ns, xc, xs + 3 * s, nrm + 3 * s, d + 2 * s, mps[depth] + b * ncs);
      }
    }
  }
  cfg->t_phase[1] = now_s() - tm;
  tm = now_s();
  for (l = depth; l >= 3; l--) {
    int Nc = cfg->order_s[l], Np = cfg->order_s[l - 1], L = Nc > Np ? Nc : Np;
    int ncc = 2 * (Nc + 1) * (Nc + 1), ncp = 2 * (Np + 1) * (Np + 1);
    double wc = t->D / (1 << l);
    double *E = calloc(2 * (size_t)((L + 2) * (L + 3) * (L + 4) / 6), sizeof(double));
    double *wk1 = malloc(sizeof(double) * 2 * (L + 1) * (L + 1));
    double *wk2 = malloc(sizeof(double) * 2 * (L + 1) * (L + 1));
    uint64_t np = 1ull << (3 * (l - 1)), p;
    orc_hel_coaxial_table(0, L, k * wc * 0.5 * sqrt(3.0), E);
    for (p = 0; p < np; p++) {
      int c;
      double *Cp = mps[l - 1] + p * ncp;
      memset(Cp, 0, sizeof(double) * ncp);
      if (t->bcount[l - 1][p] == 0) continue;
      for (c = 0; c < 8; c++) {
        double th, ph, ch;
        if (t->bcount[l][8 * p + c] == 0) continue;
        child_angles(c, 1, &th, &ph, &ch);
        orc_hel_rotate(wk1, mps[l] + (8 * p + c) * ncc, Nc, Hud[0][c], ph, ch, 0.0);
        orc_hel_coaxial(wk2, Np, wk1, Nc, E, L, 0, 0.0);
        orc_hel_rotate(Cp, wk2, Np, Hud[0][c], ch, ph, 1.0);
      }
    }
    free(E); free(wk1); free(wk2);
  }
  cfg->t_phase[2] = now_s() - tm;
  tm = now_s();
  for (l = 2; l <= depth; l++) {
    int Ns = cfg->order_s[l], Nr = cfg->order_r[l], L = Ns > Nr ? Ns : Nr, o;
    double w = t->D / (1 << l);
    double *SR[343] = {0};
    hel_job jobs[64];
    pthread_t th[64];
    int nth = cfg->nthreads > 1 ? cfg->nthreads : 1;
    if (nth > 64) nth = 64;
    for (o = 0; o < 343; o++) {
      int dx = o / 49 - 3, dy = (o / 7) % 7 - 3, dz = o % 7 - 3, o2;
      double r2;
      if (abs(dx) <= 1 && abs(dy) <= 1 && abs(dz) <= 1) continue;
      r2 = sqrt((double)(dx * dx + dy * dy + dz * dz));
      /* share the table between offsets of equal length */
      for (o2 = 0; o2 < o; o2++) {
        int ex = o2 / 49 - 3, ey = (o2 / 7) % 7 - 3, ez = o2 % 7 - 3;
        if (SR[o2] && ex * ex + ey * ey + ez * ez == dx * dx + dy * dy + dz * dz) {
          SR[o] = SR[o2];
          break;
        }
      }
      if (!SR[o]) {
        SR[o] = calloc(2 * (size_t)((L + 2) * (L + 3) * (L + 4) / 6), sizeof(double));
        orc_hel_coaxial_table(1, L, k * w * r2, SR[o]);
      }
    }
    for (i = 0; i < nth; i++) {
      jobs[i].t = t; jobs[i].cfg = cfg; jobs[i].level = l; jobs[i].nth = nth;
      jobs[i].ith = i; jobs[i].mps = mps; jobs[i].mpr = mpr; jobs[i].Htab = Htab;
      jobs[i].nrot = nrot; jobs[i].SR = SR;
    }
    if (nth == 1) hel_m2l_worker(&jobs[0]);
    else {
      for (i = 0; i < nth; i++) pthread_create(&th[i], NULL, hel_m2l_worker, &jobs[i]);
      for (i = 0; i < nth; i++) pthread_join(th[i], NULL);
    }
    for (o = 0; o < 343; o++) {
      int o2, shared = 0;
      if (!SR[o]) continue;
      for (o2 = o + 1; o2 < 343; o2++) if (SR[o2] == SR[o]) SR[o2] = NULL, shared = 1;
      (void)shared;
      free(SR[o]);
    }
    if (l < depth) {
      int Np = cfg->order_r[l], Nc = cfg->order_r[l + 1], L2 = Np > Nc ? Np : Nc;
      int ncp = 2 * (Np + 1) * (Np + 1), ncc = 2 * (Nc + 1) * (Nc + 1);
      double wc = t->D / (1 << (l + 1));
      double *E = calloc(2 * (size_t)((L2 + 2) * (L2 + 3) * (L2 + 4) / 6), sizeof(double));
      double *wk1 = malloc(sizeof(double) * 2 * (L2 + 1) * (L2 + 1));
      double *wk2 = malloc(sizeof(double) * 2 * (L2 + 1) * (L2 + 1));
      uint64_t np = 1ull << (3 * l), p;
      orc_hel_coaxial_table(0, L2, k * wc * 0.5 * sqrt(3.0), E);
      for (p = 0; p < np; p++) {
        int c;
        for (c = 0; c < 8; c++) {
          double th, ph, ch;
          child_angles(c, -1, &th, &ph, &ch);
          orc_hel_rotate(wk1, mpr[l] + p * ncp, Np, Hud[1][c], ph, ch, 0.0);
          orc_hel_coaxial(wk2, Nc, wk1, Np, E, L2, 0, 0.0);
          orc_hel_rotate(mpr[l + 1] + (8 * p + c) * ncc, wk2, Nc, Hud[1][c], ch, ph, 1.0);
        }
      }
      free(E); free(wk1); free(wk2);
    }
  }
  cfg->t_phase[3] = now_s() - tm;
  tm = now_s();
  {
    int Nr = cfg->order_r[depth], ncr = 2 * (Nr + 1) * (Nr + 1);
    for (i = 0; i < ntgt; i++) {
      const double *x = xt + 3 * i;
      uint64_t b = orc_point_index(x, t->x0, t->D) >> (3 * (20 - depth));
      uint64_t nb[27];
      double xc[3], loc[8] = {0};
      int nn, a, j, c;
      box_centre(t, depth, b, xc);
      orc_hel_j_eval(k, Nr, xc, mpr[depth] + b * ncr, x, cfg->field, loc);
      nn = orc_box_neighbours(depth, b, nb);
      for (a = 0; a < nn; a++)
        for (j = 0; j < t->bcount[depth][nb[a]]; j++) {
          int s = t->ip[t->bstart[depth][nb[a]] + j];
          double rv[3] = {x[0] - xs[3 * s], x[1] - xs[3 * s + 1], x[2] - xs[3 * s + 2]};
          double r = sqrt(rv[0] * rv[0] + rv[1] * rv[1] + rv[2] * rv[2]);
          cplx G, dG, ddG, pot = 0, g[3] = {0, 0, 0};
          if (r <= CUTOFF) continue;
          G = cexp(I * k * r) / (4.0 * M_PI * I * k * r);
          dG = G * (I * k - 1.0 / r);
          ddG = G * ((I * k - 1.0 / r) * (I * k - 1.0 / r) + 1.0 / (r * r));
          if (q) {
            cplx qq = q[2 * s] + I * q[2 * s + 1];
            pot += qq * G;
            for (c = 0; c < 3; c++) g[c] += qq * dG * rv[c] / r;
          }
          if (d) {
            cplx dd = d[2 * s] + I * d[2 * s + 1];
            const double *nv = nrm + 3 * s;
            double nr = (nv[0] * rv[0] + nv[1] * rv[1] + nv[2] * rv[2]) / r;
            pot += dd * dG * nr;
            for (c = 0; c < 3; c++)
              g[c] += dd * (ddG * nr * rv[c] / r + dG * (nv[c] - nr * rv[c] / r) / r);
          }
          loc[0] += creal(pot); loc[1] += cimag(pot);
          for (c = 0; c < 3; c++) { loc[2 + 2 * c] += creal(g[c]); loc[3 + 2 * c] += cimag(g[c]); }
        }
      if (cfg->field & 1) { f[8 * i] += loc[0]; f[8 * i + 1] += loc[1]; }
      if (cfg->field & 2) for (c = 2; c < 8; c++) f[8 * i + c] += loc[c];
    }
  }
  cfg->t_phase[4] = now_s() - tm;
  if (cfg->dump_level >= 2 && cfg->dump_level <= depth) {
    int l2 = cfg->dump_level;
    size_t nb = (size_t)1 << (3 * l2);
    if (cfg->dump_mps)
      memcpy(cfg->dump_mps, mps[l2], sizeof(double) * nb * 2 * (cfg->order_s[l2] + 1) * (cfg->order_s[l2] + 1));
    if (cfg->dump_mpr)
      memcpy(cfg->dump_mpr, mpr[l2], sizeof(double) * nb * 2 * (cfg->order_r[l2] + 1) * (cfg->order_r[l2] + 1));
  }
  for (l = 2; l <= depth; l++) { free(mps[l]); free(mpr[l]); }
  for (i = 0; i < 8; i++) { free(Hud[0][i]); free(Hud[1][i]); }
  free(Htab);
  return 0;
}

int orc_fmm(orc_fmm_cfg *cfg, int nsrc, const double *xs, const double *q,
            const double *nrm, const double *d, int ntgt, const double *xt,
            double *f)
{
  orc_tree t;
  int rc;
  double t0 = now_s(), t1;
  memset(&t, 0, sizeof(t));
  t.depth = cfg->depth;
  t.x0[0] = cfg->x0[0]; t.x0[1] = cfg->x0[1]; t.x0[2] = cfg->x0[2];
  t.D = cfg->D;
  if (cfg->depth < 2 || cfg->depth > 16) return -2;
  if (tree_build(&t, nsrc, xs) != 0) { tree_free(&t); return -1; }
  t1 = now_s();
  cfg->t_phase[0] = t1 - t0;
  if (cfg->problem == 1)
    rc = lap_solve(cfg, &t, xs, q, nrm, d, ntgt, xt, f);
  else if (cfg->problem == 2)
    rc = hel_solve(cfg, &t, xs, q, nrm, d, ntgt, xt, f);
  else
    rc = -3;
  tree_free(&t);
  cfg->t_phase[5] = now_s() - t0;
  return rc;
}
