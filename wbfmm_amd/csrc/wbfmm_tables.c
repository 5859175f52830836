/*
 * wbfmm_tables.c -- FP64 operator tables for the GPU passes, built on the
 * host and uploaded once per order.
 *
 * Rotations: the reference rotates an expansion with the Gumerov &
 * Duraiswami coefficients H_n^{nu,m}(theta) and the phases e^{-i nu phi},
 * e^{i m chi} (src/rotations.c:131-207, src/rotations-laplace.c:36-118).
 * For the GPU each rotation is turned into its real (2n+1)x(2n+1) blocks
 * acting on the real-packed coefficients, one set per translation
 * direction: the 316 list-4 offsets (M2L), the 8 child->parent (M2M) and
 * 8 parent->child (L2L) directions.  Directions sharing theta share H, as
 * in the reference's 49-entry table (src/shift.c:628-775).
 *
 * Coaxial tables: the closed forms of src/laplace.c:542-725 at unit
 * distance, c[kind][(m*(L+1)+n)*(L+1)+nu], kind 0 SS, 1 RR, 2 SR.
 */

#include <math.h>
#include <stdlib.h>
#include <string.h>

#include <quadmath.h>

#include "wbfmm_hip.h"

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

#define RIDX(nu, m, n) ((n) * (4 * (n) - 1) * ((n) + 1) / 6 + ((nu) + (n)) * ((n) + 1) + (m))

static size_t rot_size(int N) { return (size_t)RIDX(N + 1, 0, N + 1); }

/* Gumerov-Duraiswami recursion coefficients (G&D 2003, eqs 5.48-5.55).
   The recursion for H runs from degree 2N down while m goes up, and in
   FP64 it loses about 0.6 decimal digits per order (2e-10 at N=12, 7e-6
   at N=20); it is therefore evaluated in quad precision (__float128) and
   rounded to FP64 once, which keeps every table entry within a few ulp up
   to order 40. */
typedef __float128 qreal;

static qreal ga(int n, int m)
{
  m = abs(m);
  if (n < m) return 0;
  return sqrtq((qreal)(n + 1 + m) * (n + 1 - m) / ((qreal)(2 * n + 1) * (2 * n + 3)));
}

static qreal gb(int n, int m)
{
  qreal v;
  if (n < abs(m)) return 0;
  v = sqrtq((qreal)(n - m - 1) * (n - m) / ((qreal)(2 * n - 1) * (2 * n + 1)));
  return m >= 0 ? v : -v;
}

/* H_n^{nu,m}(theta) for n<=N: m=0 column from normalised Legendre
   functions up to degree 2N, then recursion 5.55 advancing m. */
static void H_rotation(double *H, int N, double th)
{
  int L = 2 * N, n, m, nu;
  size_t i, nr = rot_size(N);
  qreal C = cosq((qreal)th), S = sinq((qreal)th);
  qreal *h = calloc(rot_size(L), sizeof(qreal));
  qreal *P = malloc(sizeof(qreal) * (size_t)(L + 2) * (L + 3) / 2);
#define PI_(n, m) ((n) * ((n) + 1) / 2 + (m))
  P[0] = 1 / sqrtq(4 * M_PIq);
  if (L >= 1) {
    P[PI_(1, 0)] = C * sqrtq((qreal)3) * P[0];
    P[PI_(1, 1)] = S * sqrtq((qreal)1.5Q) * P[0];
  }
  for (n = 1; n < L; n++) {
    P[PI_(n + 1, n + 1)] = sqrtq((qreal)(2 * n + 3) / (2 * n + 2)) * S * P[PI_(n, n)];
    for (m = 0; m <= n; m++) {
      qreal pm1 = (m <= n - 1) ? P[PI_(n - 1, m)] : 0;
      P[PI_(n + 1, m)] = (C * sqrtq((qreal)(2 * n + 1)) * P[PI_(n, m)] -
                          sqrtq((qreal)(n * n - m * m) / (2 * n - 1)) * pm1) /
                         sqrtq((qreal)((n + 1) * (n + 1) - m * m) / (2 * n + 3));
    }
  }
  for (n = 0; n <= L; n++)
    for (nu = 0; nu <= n; nu++) {
      qreal v = P[PI_(n, nu)] * sqrtq(4 * M_PIq / (2 * n + 1));
      h[RIDX(nu, 0, n)] = v;
      h[RIDX(-nu, 0, n)] = v;
    }
#undef PI_
  for (m = 0; m <= N; m++)
    for (n = m + 2; n <= L - m; n++) {
      qreal b = gb(n, m);
      for (nu = -n + 1; nu <= n - 1; nu++)
        h[RIDX(nu, m + 1, n - 1)] =
            ((gb(n, -nu - 1) * (1 - C) * h[RIDX(nu + 1, m, n)] -
              gb(n, nu - 1) * (1 + C) * h[RIDX(nu - 1, m, n)]) / 2 -
             ga(n - 1, nu) * S * h[RIDX(nu, m, n)]) / b;
    }
  for (i = 0; i < nr; i++) H[i] = (double)h[i];
  free(h);
  free(P);
}

/* real blocks of Co_n^nu = e^{-i nu ph} sum_m H_n^{nu,m} e^{i m ch} Ci_n^m
   on real-packed coefficients; R[n] row-major (2n+1)^2 at
   n(2n-1)(2n+1)/3 */
static void real_blocks(double *R, const double *H, int N, double ph,
                        double ch)
{
  int n, nu, m;
  for (n = 0; n <= N; n++) {
    int w = 2 * n + 1;
    double *B = R + (size_t)n * (2 * n - 1) * (2 * n + 1) / 3;
    memset(B, 0, sizeof(double) * w * w);
    for (nu = 0; nu <= n; nu++) {
      /* coefficients of tr and ti with respect to the inputs */
      double dtr[2 * 64 + 1] = {0}, dti[2 * 64 + 1] = {0};
      double hp0 = H[RIDX(nu, 0, n)];
      dtr[0] = hp0;
      for (m = 1; m <= n; m++) {
        double hp = H[RIDX(nu, m, n)], hm = H[RIDX(-nu, m, n)];
        double cm = cos(m * ch), sm = sin(m * ch);
        dtr[2 * m - 1] = (hp + hm) * cm;
        dtr[2 * m] = -(hp + hm) * sm;
        dti[2 * m - 1] = (hp - hm) * sm;
        dti[2 * m] = (hp - hm) * cm;
      }
      if (nu == 0) {
        for (m = 0; m < w; m++) B[0 * w + m] = dtr[m];
      } else {
        double cn = cos(nu * ph), sn = sin(nu * ph);
        for (m = 0; m < w; m++) {
          B[(2 * nu - 1) * w + m] = cn * dtr[m] + sn * dti[m];
          B[(2 * nu) * w + m] = cn * dti[m] - sn * dtr[m];
        }
      }
    }
  }
}

/* angles of a translation along e (not normalised): theta polar angle;
   phi = -pi/2 and chi = azimuth, or both 0 on the z axis (the frame of
   src/util.c:224-262 with the angles of src/rotations.c:99-128) */
static void dir_angles(double ex, double ey, double ez, double *th,
                       double *ph, double *ch)
{
  double r = sqrt(ex * ex + ey * ey + ez * ez);
  *th = acos(ez / r);
  if (ex == 0.0 && ey == 0.0) {
    *ph = 0.0;
    *ch = 0.0;
  } else {
    *ph = -0.5 * M_PI;
    *ch = atan2(ey, ex);
  }
}

/* H(theta) tables are shared by every direction with the same polar
   angle (49 distinct values over the list-4 offsets, as in the reference's
   table, src/shift.c:637-686); cache them while building */
typedef struct {
  int n, L;
  double th[400];
  double *H[400];
} hcache;

static const double *cached_H(hcache *hc, double th)
{
  int i;
  for (i = 0; i < hc->n; i++)
    if (hc->th[i] == th) return hc->H[i];
  hc->th[hc->n] = th;
  hc->H[hc->n] = malloc(sizeof(double) * rot_size(hc->L));
  H_rotation(hc->H[hc->n], hc->L, th);
  return hc->H[hc->n++];
}

static void add_direction(double *host, size_t fwd, size_t back, hcache *hc,
                          double ex, double ey, double ez)
{
  double th, ph, ch;
  const double *H;
  dir_angles(ex, ey, ez, &th, &ph, &ch);
  H = cached_H(hc, th);
  real_blocks(host + fwd, H, hc->L, ph, ch);
  real_blocks(host + back, H, hc->L, ch, ph);
}

int wbfmm_hip_rotation_tables(int L, double **dev, wbfmm_rot_layout_t *lay)
{
  size_t nrot, total;
  double *host;
  hcache hc;
  int dx, dy, dz, c, rc, i;
  if (L < 0 || L > 64) return WBFMM_HIP_ERR_ORDER;
  nrot = (size_t)(L + 1) * (2 * L + 1) * (2 * L + 3) / 3;
  memset(lay, 0, sizeof(*lay));
  lay->L = L;
  lay->nrot = nrot;
  lay->m2l_fwd = 0;
  lay->m2l_back = 343 * nrot;
  lay->up_fwd = 2 * 343 * nrot;
  lay->up_back = lay->up_fwd + 8 * nrot;
  lay->dn_fwd = lay->up_back + 8 * nrot;
  lay->dn_back = lay->dn_fwd + 8 * nrot;
  lay->total = lay->dn_back + 8 * nrot;
  total = lay->total + 343; /* offset lengths appended */
  host = calloc(total, sizeof(double));
  if (!host) return WBFMM_HIP_ERR_NOMEM;
  hc.n = 0;
  hc.L = L;
  for (dx = -3; dx <= 3; dx++)
    for (dy = -3; dy <= 3; dy++)
      for (dz = -3; dz <= 3; dz++) {
        int o = (dx + 3) * 49 + (dy + 3) * 7 + dz + 3;
        if (abs(dx) <= 1 && abs(dy) <= 1 && abs(dz) <= 1) continue;
        /* translation from source to target: e = target - source */
        add_direction(host, lay->m2l_fwd + o * nrot, lay->m2l_back + o * nrot,
                      &hc, -dx, -dy, -dz);
        lay->r[o] = sqrt((double)(dx * dx + dy * dy + dz * dz));
        host[lay->total + o] = lay->r[o];
      }
  for (c = 0; c < 8; c++) {
    double ex = 0.5 - (c & 1), ey = 0.5 - ((c >> 1) & 1), ez = 0.5 - ((c >> 2) & 1);
    add_direction(host, lay->up_fwd + c * nrot, lay->up_back + c * nrot, &hc,
                  ex, ey, ez);
    add_direction(host, lay->dn_fwd + c * nrot, lay->dn_back + c * nrot, &hc,
                  -ex, -ey, -ez);
  }
  for (i = 0; i < hc.n; i++) free(hc.H[i]);
  rc = wbfmm_hip_malloc((void **)dev, sizeof(double) * total);
  if (rc == 0) rc = wbfmm_hip_memcpy(*dev, host, sizeof(double) * total);
  free(host);
  return rc;
}

static double lap_cfft(int kind, int n, int nu, int m)
{
  double c;
  if (kind == 0) {
    c = lgamma(n - m + 1) + lgamma(n + m + 1) - lgamma(nu - m + 1) -
        lgamma(nu + m + 1) - 2.0 * lgamma(n - nu + 1);
    return exp(0.5 * (c + log((2.0 * nu + 1) / (2.0 * n + 1))));
  }
  if (kind == 1) {
    c = lgamma(nu - m + 1) + lgamma(nu + m + 1) - lgamma(n - m + 1) -
        lgamma(n + m + 1) - 2.0 * lgamma(nu - n + 1);
    return exp(0.5 * (c + log((2.0 * nu + 1) / (2.0 * n + 1))));
  }
  c = 2.0 * lgamma(n + nu + 1) - lgamma(n - m + 1) - lgamma(n + m + 1) -
      lgamma(nu - m + 1) - lgamma(nu + m + 1);
  c = exp(0.5 * (c + log((2.0 * nu + 1) / (2.0 * n + 1))));
  return ((n + m) % 2) ? -c : c;
}

int wbfmm_hip_laplace_coaxial_tables(int L, double **dev)
{
  size_t L1 = (size_t)L + 1, per = L1 * L1 * L1;
  double *host = calloc(3 * per, sizeof(double));
  int kind, m, n, nu, rc;
  if (!host) return WBFMM_HIP_ERR_NOMEM;
  for (kind = 0; kind < 3; kind++)
    for (m = 0; m <= L; m++)
      for (n = m; n <= L; n++)
        for (nu = m; nu <= L; nu++) {
          if (kind == 0 && nu > n) continue;
          if (kind == 1 && nu < n) continue;
          host[kind * per + ((size_t)m * L1 + n) * L1 + nu] = lap_cfft(kind, n, nu, m);
        }
  rc = wbfmm_hip_malloc((void **)dev, sizeof(double) * 3 * per);
  if (rc == 0) rc = wbfmm_hip_memcpy(*dev, host, sizeof(double) * 3 * per);
  free(host);
  return rc;
}

/* ------------------------------------------------------------------
 * MFMA form of the Laplace M2L (v_mfma_f64_16x16x4f64 operands).
 *
 * The three factors of one translation (rotate, coaxial S|R, rotate
 * back; src/pass-laplace.c:73-101) are applied to 16 boxes at a time as
 * 16x16x4 FP64 matrix products.  Rows are padded to 16-row tiles; the
 * inputs of a tile span whole degree blocks (rotations) or order blocks
 * (coaxial), so each tile needs a short run of 4-wide k steps.  The A
 * operand of each (tile, k step) is stored in lane order: lane l holds
 * A[l % 16][l / 16].  The coaxial factor is stored at unit distance; the
 * kernel scales its inputs by t^-nu and outputs by t^-(n+1).
 * ------------------------------------------------------------------ */

static void tile_ranges(int C, int NT, const int *blk_lo, const int *blk_hi,
                        int *kb, int *nk, int *off, int *total)
{
  int i, acc = 0;
  for (i = 0; i < NT; i++) {
    int r0 = 16 * i, r1 = 16 * i + 15, lo, hi;
    if (r1 >= C) r1 = C - 1;
    lo = blk_lo[r0];
    hi = blk_hi[r1];
    kb[i] = lo;
    nk[i] = (hi - lo + 3) / 4;
    off[i] = acc;
    acc += nk[i];
  }
  *total = acc;
}

/* store rows [16i,16i+16) x cols [kb+4j, kb+4j+4) of the dense C x C
   matrix M (row-major, zero outside) in MFMA lane order */
static void swizzle(double *dst, const double *M, int C, int NT,
                    const int *kb, const int *nk)
{
  int i, j, l;
  for (i = 0; i < NT; i++)
    for (j = 0; j < nk[i]; j++)
      for (l = 0; l < 64; l++) {
        int r = 16 * i + (l % 16), c = kb[i] + 4 * j + (l / 16);
        *dst++ = (r < C && c < C) ? M[(size_t)r * C + c] : 0.0;
      }
}

int wbfmm_hip_lap_m2l_plan(int N, wbfmm_m2l_plan_t *P, void **plan_dev,
                           double **A1, double **A3, double **A2)
{
  int C = (N + 1) * (N + 1), NT = (C + 15) / 16, n, m, r, i, dx, dy, dz, rc;
  int lo_n[WBFMM_M2L_MAXC], hi_n[WBFMM_M2L_MAXC];
  int lo_m[WBFMM_M2L_MAXC], hi_m[WBFMM_M2L_MAXC];
  double *M, *h1, *h3, *h2, *R, *Rb, *H;
  hcache hc;
  size_t per;
  if (C > WBFMM_M2L_MAXC || NT > WBFMM_M2L_MAXT) return WBFMM_HIP_ERR_ORDER;
  memset(P, 0, sizeof(*P));
  P->N = N;
  P->C = C;
  P->NT = NT;
  P->CP = 16 * NT;
  /* n-major rows: degree blocks [n^2, (n+1)^2) */
  for (n = 0; n <= N; n++)
    for (r = n * n; r < (n + 1) * (n + 1); r++) {
      lo_n[r] = n * n;
      hi_n[r] = (n + 1) * (n + 1);
      P->degn[r] = n;
    }
  /* m-major rows: m=0 (n=0..N), then for m>=1 re (n=m..N), im (n=m..N) */
  r = 0;
  for (m = 0; m <= N; m++) {
    int part, nparts = (m == 0) ? 1 : 2;
    for (part = 0; part < nparts; part++) {
      int start = r, len = N + 1 - m;
      for (n = m; n <= N; n++, r++) {
        int nrow = (m == 0) ? n * n : n * n + 2 * m - 1 + part;
        P->npos[r] = nrow;
        P->mpos[nrow] = r;
        P->degm[r] = n;
        lo_m[r] = start;
        hi_m[r] = start + len;
      }
    }
  }
  tile_ranges(C, NT, lo_n, hi_n, P->kb1, P->nk1, P->off1, &P->T13);
  tile_ranges(C, NT, lo_m, hi_m, P->kb2, P->nk2, P->off2, &P->T2);

  /* stage 2: unit-distance S|R in m-major order */
  M = calloc((size_t)C * C, sizeof(double));
  for (r = 0; r < C; r++) {
    int n1 = P->degm[r], c;
    for (c = lo_m[r]; c < hi_m[r]; c++) {
      int nu = P->degm[c], mm;
      /* order m of this block: rows of one block share m */
      int nrow = P->npos[r];
      mm = (nrow == n1 * n1) ? 0 : (nrow - n1 * n1 + 1) / 2;
      M[(size_t)r * C + c] = lap_cfft(2, n1, nu, mm);
    }
  }
  h2 = malloc(sizeof(double) * (size_t)P->T2 * 64);
  swizzle(h2, M, C, NT, P->kb2, P->nk2);

  /* stages 1 and 3: rotation blocks of every list-4 offset */
  per = (size_t)P->T13 * 64;
  h1 = calloc(343 * per, sizeof(double));
  h3 = calloc(343 * per, sizeof(double));
  R = malloc(sizeof(double) * (size_t)(N + 1) * (2 * N + 1) * (2 * N + 3) / 3);
  Rb = malloc(sizeof(double) * (size_t)(N + 1) * (2 * N + 1) * (2 * N + 3) / 3);
  hc.n = 0;
  hc.L = N;
  for (dx = -3; dx <= 3; dx++)
    for (dy = -3; dy <= 3; dy++)
      for (dz = -3; dz <= 3; dz++) {
        int o = (dx + 3) * 49 + (dy + 3) * 7 + dz + 3;
        double th, ph, ch;
        if (abs(dx) <= 1 && abs(dy) <= 1 && abs(dz) <= 1) continue;
        P->r[o] = sqrt((double)(dx * dx + dy * dy + dz * dz));
        dir_angles(-dx, -dy, -dz, &th, &ph, &ch);
        H = (double *)cached_H(&hc, th);
        real_blocks(R, H, N, ph, ch);
        real_blocks(Rb, H, N, ch, ph);
        /* expand the degree blocks into dense C x C (row-major) */
        memset(M, 0, sizeof(double) * (size_t)C * C);
        for (n = 0; n <= N; n++) {
          int w = 2 * n + 1, a, b;
          const double *B = R + (size_t)n * (2 * n - 1) * (2 * n + 1) / 3;
          for (a = 0; a < w; a++)
            for (b = 0; b < w; b++) M[(size_t)(n * n + a) * C + n * n + b] = B[a * w + b];
        }
        swizzle(h1 + o * per, M, C, NT, P->kb1, P->nk1);
        memset(M, 0, sizeof(double) * (size_t)C * C);
        for (n = 0; n <= N; n++) {
          int w = 2 * n + 1, a, b;
          const double *B = Rb + (size_t)n * (2 * n - 1) * (2 * n + 1) / 3;
          for (a = 0; a < w; a++)
            for (b = 0; b < w; b++) M[(size_t)(n * n + a) * C + n * n + b] = B[a * w + b];
        }
        swizzle(h3 + o * per, M, C, NT, P->kb1, P->nk1);
      }
  for (i = 0; i < hc.n; i++) free(hc.H[i]);
  free(R);
  free(Rb);
  free(M);
  rc = wbfmm_hip_malloc((void **)A1, sizeof(double) * 343 * per);
  if (!rc) rc = wbfmm_hip_memcpy(*A1, h1, sizeof(double) * 343 * per);
  if (!rc) rc = wbfmm_hip_malloc((void **)A3, sizeof(double) * 343 * per);
  if (!rc) rc = wbfmm_hip_memcpy(*A3, h3, sizeof(double) * 343 * per);
  if (!rc) rc = wbfmm_hip_malloc((void **)A2, sizeof(double) * (size_t)P->T2 * 64);
  if (!rc) rc = wbfmm_hip_memcpy(*A2, h2, sizeof(double) * (size_t)P->T2 * 64);
  if (!rc) rc = wbfmm_hip_malloc(plan_dev, sizeof(*P));
  if (!rc) rc = wbfmm_hip_memcpy(*plan_dev, P, sizeof(*P));
  free(h1);
  free(h2);
  free(h3);
  return rc;
}
