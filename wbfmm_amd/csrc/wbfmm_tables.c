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

#include <pthread.h>
#include <quadmath.h>
#include <unistd.h>

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
/* n independent jobs fn(i, ctx) on up to 16 host threads (the tables of
   different angles or distances do not share state) */
typedef struct {
  void (*fn)(int, void *);
  void *ctx;
  int n, stride, first;
} par_job;

static void *par_run(void *a)
{
  par_job *j = (par_job *)a;
  int i;
  for (i = j->first; i < j->n; i += j->stride) j->fn(i, j->ctx);
  return NULL;
}

static void par_for(int n, void (*fn)(int, void *), void *ctx)
{
  long ncpu = sysconf(_SC_NPROCESSORS_ONLN);
  int t, nt = n < 16 ? n : 16, started = 0;
  pthread_t th[16];
  par_job jobs[16];
  if (ncpu > 0 && nt > ncpu) nt = (int)ncpu;
  if (nt <= 1) {
    for (t = 0; t < n; t++) fn(t, ctx);
    return;
  }
  for (t = 0; t < nt; t++) {
    jobs[t].fn = fn;
    jobs[t].ctx = ctx;
    jobs[t].n = n;
    jobs[t].stride = nt;
    jobs[t].first = t;
  }
  for (t = 1; t < nt; t++)
    if (pthread_create(&th[t], NULL, par_run, &jobs[t]) == 0) started |= 1 << t;
    else par_run(&jobs[t]);  /* no thread: run the share here */
  par_run(&jobs[0]);
  for (t = 1; t < nt; t++)
    if (started & (1 << t)) pthread_join(th[t], NULL);
}

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

/* the H tables of every polar angle the list-4 offsets and the 8 child
   directions use, computed on host threads into the cache */
typedef struct {
  hcache *hc;
} hfill_job;

static void hfill_one(int i, void *c)
{
  hcache *hc = ((hfill_job *)c)->hc;
  H_rotation(hc->H[i], hc->L, hc->th[i]);
}

static void prefill_H(hcache *hc)
{
  int dx, dy, dz, c, i, k;
  double e[8 + 343][3];
  hfill_job j;
  k = 0;
  for (dx = -3; dx <= 3; dx++)
    for (dy = -3; dy <= 3; dy++)
      for (dz = -3; dz <= 3; dz++) {
        if (abs(dx) <= 1 && abs(dy) <= 1 && abs(dz) <= 1) continue;
        e[k][0] = -dx;
        e[k][1] = -dy;
        e[k][2] = -dz;
        k++;
      }
  for (c = 0; c < 8; c++) {
    e[k][0] = 0.5 - (c & 1);
    e[k][1] = 0.5 - ((c >> 1) & 1);
    e[k][2] = 0.5 - ((c >> 2) & 1);
    k++;
  }
  hc->n = 0;
  for (i = 0; i < k; i++) {
    double th, ph, ch;
    int q, s;
    for (s = 1; s >= -1; s -= 2) {
      dir_angles(s * e[i][0], s * e[i][1], s * e[i][2], &th, &ph, &ch);
      for (q = 0; q < hc->n; q++)
        if (hc->th[q] == th) break;
      if (q == hc->n && hc->n < 400) {
        double *H = malloc(sizeof(double) * rot_size(hc->L));
        if (!H) continue; /* left to cached_H */
        hc->th[hc->n] = th;
        hc->H[hc->n] = H;
        hc->n++;
      }
    }
  }
  j.hc = hc;
  par_for(hc->n, hfill_one, &j);
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
  prefill_H(&hc);
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
 * One translation of the reference (rotate, coaxial S|R, rotate back;
 * src/pass-laplace.c:73-101, src/rotations-laplace.c:36-216,
 * src/laplace.c:651-696) is factored as
 *   L += Zout(chi) H3(theta) T(t) H1(theta) Zin(chi) M
 * where Zin/Zout are the azimuthal phases (applied by the kernel on the
 * vector units), H1 = real blocks of e^{-i nu phi} H(theta) with
 * phi = -pi/2 and H3 those of H(theta) e^{i m phi'} with phi' = -pi/2: the
 * inverse rotation swaps the two azimuths, so the pair is exact for any
 * choice of phi.  With phi = -pi/2 the real (2n+1)^2 block of degree n
 * maps the real parts (m = 0..n) and the imaginary parts (m = 1..n)
 * separately: a cos block of size n+1 and a sin block of size n.  T is the
 * unit-distance coaxial S|R shift (one block per order m, shared by the
 * two parts); the kernel scales its inputs by t^-n and outputs by
 * t^-(n+1).  Blocks are bin-packed into 16-row MFMA tiles.
 * ------------------------------------------------------------------ */

/* first-fit decreasing packing of nblk blocks (sizes sz) into bins of 16
   rows; ties in size keep block order.  bin[i] gets the bin of block i,
   pos[i] its first slot within the bin; returns the number of bins and
   fills fill[b] (rows used by bin b). */
/* Blocks sorted by size (stable, decreasing).  A block of more than 16
   rows takes ceil(size/16) consecutive bins of its own (their fill set to
   16 so nothing else joins them); the others go first fit into shared
   bins.  kb[b]: first input row of bin b's k-steps, nk[b]: their number. */
static int pack_bins(int nblk, const int *sz, int *bin, int *pos, int *fill, int *kb, int *nk)
{
  int order[2 * WBFMM_M2L_MAXN + 2], i, j, b, nb = 0;
  for (i = 0; i < nblk; i++) order[i] = i;
  for (i = 1; i < nblk; i++) {        /* stable insertion sort, size desc */
    int v = order[i];
    for (j = i; j > 0 && sz[order[j - 1]] < sz[v]; j--) order[j] = order[j - 1];
    order[j] = v;
  }
  for (i = 0; i < nblk; i++) {
    int k = order[i];
    if (sz[k] > 16) {
      int nt = (sz[k] + 15) / 16, t;
      if (nb + nt > WBFMM_M2L_MAXB) return -1;
      for (t = 0; t < nt; t++) {
        fill[nb + t] = 16;
        kb[nb + t] = 16 * nb;
        nk[nb + t] = (sz[k] + 3) / 4;
      }
      bin[k] = nb;
      pos[k] = 0;
      nb += nt;
      continue;
    }
    for (b = 0; b < nb; b++)
      if (fill[b] + sz[k] <= 16) break;
    if (b == nb) {
      if (nb == WBFMM_M2L_MAXB) return -1;
      fill[nb] = 0;
      kb[nb] = -1;   /* shared bin */
      nb++;
    }
    bin[k] = b;
    pos[k] = fill[b];
    fill[b] += sz[k];
  }
  for (b = 0; b < nb; b++)
    if (kb[b] < 0) {
      kb[b] = 16 * b;
      nk[b] = (fill[b] + 3) / 4;
    }
  return nb;
}

/* blocks of stages 1 and 3: 0..N cos part of degree n (size n+1),
   N+1..2N sin part of degree n = i-N (size n); stage 2: 0..N order m
   cos part (size N+1-m), N+1..2N order m = i-N sin part */
static int blk_size13(int N, int i) { return i <= N ? i + 1 : i - N; }
static int blk_size2(int N, int i) { return i <= N ? N + 1 - i : N + 1 - (i - N); }

typedef struct {
  int nb, bin[2 * WBFMM_M2L_MAXN + 2], pos[2 * WBFMM_M2L_MAXN + 2];
  int fill[WBFMM_M2L_MAXB], kb[WBFMM_M2L_MAXB], nk[WBFMM_M2L_MAXB];
} packing;

static int do_pack(int N, int stage2, packing *pk)
{
  int sz[2 * WBFMM_M2L_MAXN + 2], i, nblk = 2 * N + 1;
  for (i = 0; i < nblk; i++) sz[i] = stage2 ? blk_size2(N, i) : blk_size13(N, i);
  memset(pk->fill, 0, sizeof(pk->fill));
  pk->nb = pack_bins(nblk, sz, pk->bin, pk->pos, pk->fill, pk->kb, pk->nk);
  return pk->nb;
}

/* slot of entity (n, nu, part) (part 0 cos, 1 sin) as an input (k) or
   output (row) of each stage; blocks hold their entities in order */
static int slot13(const packing *p, int N, int n, int idx, int part)
{
  int blk = part ? N + n : n;             /* sin part of degree n: N + n */
  return 16 * p->bin[blk] + p->pos[blk] + (part ? idx - 1 : idx);
}
static int slot2(const packing *p, int N, int n, int m, int part)
{
  int blk = part ? N + m : m;
  return 16 * p->bin[blk] + p->pos[blk] + (n - m);
}

/* row of the cos/sin part of degree n block of real_blocks() output
   (stage 1, phi = -pi/2): cos rows re(nu) for even nu, im(nu) odd nu */
static int h1_row(int nu, int part)
{
  if (!part) return nu == 0 ? 0 : (nu % 2 == 0 ? 2 * nu - 1 : 2 * nu);
  return nu % 2 == 0 ? 2 * nu : 2 * nu - 1;
}
static int h1_col(int m, int part) { return part ? 2 * m : (m == 0 ? 0 : 2 * m - 1); }
/* stage 3 (phi' = -pi/2 on the input side): rows tr/ti, cols cos inputs
   re(m) for even m, im(m) odd m; sin inputs im(m) even, re(m) odd */
static int h3_row(int nu, int part) { return part ? 2 * nu : (nu == 0 ? 0 : 2 * nu - 1); }
static int h3_col(int m, int part)
{
  if (!part) return m == 0 ? 0 : (m % 2 == 0 ? 2 * m - 1 : 2 * m);
  return m % 2 == 0 ? 2 * m : 2 * m - 1;
}

/* store the 16 x 4nk tile of bin b from the dense slot matrix T
   (16*NB x 16*NB, row-major) in lane order */
static void put_tiles(double *dst, const double *T, const wbfmm_m2l_plan_t *P)
{
  int b, j, l, S = 16 * P->NB;
  for (b = 0; b < P->NB; b++)
    for (j = 0; j < P->nk[b]; j++)
      for (l = 0; l < 64; l++) {
        int r = 16 * b + l % 16, c = P->kbase[b] + 4 * j + l / 16;
        dst[((size_t)P->koff[b] + j) * 64 + l] = T[(size_t)r * S + c];
      }
}

static void theta_tiles(const wbfmm_m2l_plan_t *P, const packing *p13,
                        const double *H, double *A1, double *A3, double *T);

/* the rotation tiles of one polar angle (a host thread each) */
typedef struct {
  const wbfmm_m2l_plan_t *P;
  const packing *p13;
  const double *ths;
  double *A1, *A3;
  int rc;
} lap_tile_job;

static void lap_tile_one(int k, void *c)
{
  lap_tile_job *j = (lap_tile_job *)c;
  const wbfmm_m2l_plan_t *P = j->P;
  size_t per = (size_t)P->KT * 64, S = (size_t)16 * P->NB;
  double *H = malloc(sizeof(double) * rot_size(P->N)), *T = calloc(S * S, sizeof(double));
  if (!H || !T) {
    j->rc = WBFMM_HIP_ERR_NOMEM;
  } else {
    H_rotation(H, P->N, j->ths[k]);
    theta_tiles(P, j->p13, H, j->A1 ? j->A1 + k * per : NULL, j->A3 ? j->A3 + k * per : NULL,
                T);
  }
  free(H);
  free(T);
}

/* stage-1 and stage-3 tiles of one polar angle (H: its H table to order
   N); T: scratch of (16 NB)^2 doubles */
static void theta_tiles(const wbfmm_m2l_plan_t *P, const packing *p13,
                        const double *H, double *A1, double *A3, double *T)
{
  int N = P->N, S = 16 * P->NB, n, nu, m, part;
  size_t nr = (size_t)(N + 1) * (2 * N + 1) * (2 * N + 3) / 3;
  double *R = malloc(sizeof(double) * nr);
  if (A1) {
    real_blocks(R, H, N, -0.5 * M_PI, 0.0);
    memset(T, 0, sizeof(double) * (size_t)S * S);
    for (n = 0; n <= N; n++) {
      const double *B = R + (size_t)n * (2 * n - 1) * (2 * n + 1) / 3;
      int w = 2 * n + 1;
      for (part = 0; part < 2; part++)
        for (nu = part; nu <= n; nu++)
          for (m = part; m <= n; m++)
            T[(size_t)slot13(p13, N, n, nu, part) * S + slot13(p13, N, n, m, part)] =
                B[h1_row(nu, part) * w + h1_col(m, part)];
    }
    put_tiles(A1, T, P);
  }
  if (A3) {
    real_blocks(R, H, N, 0.0, -0.5 * M_PI);
    memset(T, 0, sizeof(double) * (size_t)S * S);
    for (n = 0; n <= N; n++) {
      const double *B = R + (size_t)n * (2 * n - 1) * (2 * n + 1) / 3;
      int w = 2 * n + 1;
      for (part = 0; part < 2; part++)
        for (nu = part; nu <= n; nu++)
          for (m = part; m <= n; m++)
            T[(size_t)slot13(p13, N, n, nu, part) * S + slot13(p13, N, n, m, part)] =
                B[h3_row(nu, part) * w + h3_col(m, part)];
    }
    put_tiles(A3, T, P);
  }
  free(R);
}

int wbfmm_hip_lap_m2l_tables(int N, wbfmm_m2l_plan_t *P, double *A1,
                             double *A3, double *CS)
{
  packing p13, p2;
  int b, n, nu, m, part, i, dx, dy, dz, S;
  double *T, ths[343];
  hcache hc;
  if (N < 1 || N > 27) return WBFMM_HIP_ERR_ORDER;
  memset(P, 0, sizeof(*P));
  P->N = N;
  P->C = (N + 1) * (N + 1);
  P->NU = (N + 1) * (N + 2) / 2;
  if (do_pack(N, 0, &p13) < 0 || do_pack(N, 1, &p2) < 0) return WBFMM_HIP_ERR_ORDER;
  /* the two block multisets are equal, so are the bins */
  if (p13.nb != p2.nb || 2 * (N + 1) * (N + 2) / 2 > WBFMM_M2L_MAXU * 2) return WBFMM_HIP_ERR_ORDER;
  for (b = 0; b < p13.nb; b++)
    if (p13.fill[b] != p2.fill[b] || p13.kb[b] != p2.kb[b] || p13.nk[b] != p2.nk[b])
      return WBFMM_HIP_ERR_ORDER;
  P->NB = p13.nb;
  for (b = 0; b < P->NB; b++) {
    P->nk[b] = p13.nk[b];
    P->kbase[b] = p13.kb[b];
    P->koff[b] = P->KT;
    P->KT += P->nk[b];
  }
  S = 16 * P->NB;
  P->S = S > 2 * P->NU ? S : 2 * P->NU;
  for (i = 0; i < S; i++) P->map1[i] = P->map2[i] = P->map3[i] = P->S << 6;
  for (n = 0; n <= N; n++)
    for (part = 0; part < 2; part++)
      for (nu = part; nu <= n; nu++) {
        /* stage-1 output (n,nu,part) -> stage-2 input, scale t^-n */
        P->map1[slot13(&p13, N, n, nu, part)] = (slot2(&p2, N, n, nu, part) << 6) | n;
        /* stage-2 output -> stage-3 input, scale t^-(n+1) */
        P->map2[slot2(&p2, N, n, nu, part)] = (slot13(&p13, N, n, nu, part) << 6) | (n + 1);
        /* stage-3 output tr/ti(n,nu) -> slot u or NU + u */
        P->map3[slot13(&p13, N, n, nu, part)] =
            ((part ? P->NU : 0) + n * (n + 1) / 2 + nu) << 6;
      }
  for (n = 0; n <= N; n++)
    for (m = 0; m <= n; m++)
      P->s1[n * (n + 1) / 2 + m] =
          slot13(&p13, N, n, m, 0) | ((m ? slot13(&p13, N, n, m, 1) : P->S) << 16);

  /* theta of each list-4 offset; e = target - source = -(dx,dy,dz) */
  P->NTH = 0;
  for (dx = -3; dx <= 3; dx++)
    for (dy = -3; dy <= 3; dy++)
      for (dz = -3; dz <= 3; dz++) {
        int o = (dx + 3) * 49 + (dy + 3) * 7 + dz + 3, k;
        double th, ph, ch;
        if (abs(dx) <= 1 && abs(dy) <= 1 && abs(dz) <= 1) continue;
        P->r[o] = sqrt((double)(dx * dx + dy * dy + dz * dz));
        dir_angles(-dx, -dy, -dz, &th, &ph, &ch);
        P->chi[o] = (dx == 0 && dy == 0) ? 0.0 : ch;
        for (k = 0; k < P->NR; k++)
          if (P->rr[k] == P->r[o]) break;
        if (k == P->NR) P->rr[P->NR++] = P->r[o];
        P->rix[o] = k;
        /* collinear offsets, e.g. (2,2,2) and (3,3,3), share theta up to
           rounding: 49 distinct angles as in src/shift.c:637-686 */
        for (k = 0; k < P->NTH; k++)
          if (fabs(ths[k] - th) < 1e-12) break;
        if (k == P->NTH) ths[P->NTH++] = th;
        P->th[o] = k;
      }
  if (CS)
    for (i = 0; i < 343; i++)
      for (m = 0; m <= N; m++) {
        CS[((size_t)i * (N + 1) + m) * 2] = cos(m * P->chi[i]);
        CS[((size_t)i * (N + 1) + m) * 2 + 1] = sin(m * P->chi[i]);
      }
  if (!A1 && !A3) return 0;

  {
    lap_tile_job j = {P, &p13, ths, A1, A3, 0};
    (void)hc;
    (void)T;
    (void)S;
    par_for(P->NTH, lap_tile_one, &j);
    return j.rc;
  }
}

/* stage 2 at box width w: S|R(n, nu, m) / t^(n+nu+1), t = rr[i] w, the
   same block for the cos and sin parts (src/laplace.c:651-696) */
int wbfmm_hip_lap_m2l_coax(const wbfmm_m2l_plan_t *P, double w, double *A2)
{
  packing p2;
  int N = P->N, S = 16 * P->NB, i, m, n, nu, part;
  double *T;
  if (w <= 0.0) return WBFMM_HIP_ERR_ARG;
  do_pack(N, 1, &p2);
  T = calloc((size_t)S * S, sizeof(double));
  if (!T) return WBFMM_HIP_ERR_NOMEM;
  for (i = 0; i < P->NR; i++) {
    double t = P->rr[i] * w;
    memset(T, 0, sizeof(double) * (size_t)S * S);
    for (m = 0; m <= N; m++)
      for (part = 0; part < (m ? 2 : 1); part++)
        for (n = m; n <= N; n++)
          for (nu = m; nu <= N; nu++)
            T[(size_t)slot2(&p2, N, n, m, part) * S + slot2(&p2, N, nu, m, part)] =
                lap_cfft(2, n, nu, m) * pow(t, -(double)(n + nu + 1));
    put_tiles(A2 + (size_t)i * P->KT * 64, T, P);
  }
  free(T);
  return 0;
}

int wbfmm_hip_lap_m2l_plan(int N, wbfmm_m2l_plan_t *P, void **plan_dev,
                           double **A1, double **A3, double **CS)
{
  double *h1 = NULL, *h3 = NULL, *hc;
  size_t n13, ncs;
  const int tiles = A1 != NULL && A3 != NULL;
  int rc = wbfmm_hip_lap_m2l_tables(N, P, NULL, NULL, NULL);
  if (rc) return rc;
  n13 = (size_t)P->NTH * P->KT * 64;
  ncs = (size_t)343 * (N + 1) * 2;
  if (tiles) {
    h1 = malloc(sizeof(double) * n13);
    h3 = malloc(sizeof(double) * n13);
  }
  hc = malloc(sizeof(double) * ncs);
  if ((tiles && (!h1 || !h3)) || !hc) rc = WBFMM_HIP_ERR_NOMEM;
  if (!rc) rc = wbfmm_hip_lap_m2l_tables(N, P, h1, h3, hc);
  if (!rc && tiles) rc = wbfmm_hip_malloc((void **)A1, sizeof(double) * n13);
  if (!rc && tiles) rc = wbfmm_hip_memcpy(*A1, h1, sizeof(double) * n13);
  if (!rc && tiles) rc = wbfmm_hip_malloc((void **)A3, sizeof(double) * n13);
  if (!rc && tiles) rc = wbfmm_hip_memcpy(*A3, h3, sizeof(double) * n13);
  if (!rc) rc = wbfmm_hip_malloc((void **)CS, sizeof(double) * ncs);
  if (!rc) rc = wbfmm_hip_memcpy(*CS, hc, sizeof(double) * ncs);
  if (!rc) rc = wbfmm_hip_malloc(plan_dev, sizeof(*P));
  if (!rc) rc = wbfmm_hip_memcpy(*plan_dev, P, sizeof(*P));
  free(h1);
  free(h3);
  free(hc);
  return rc;
}

/* ------------------------------------------------------------------
 * Helmholtz tables.
 * ------------------------------------------------------------------ */

/* geometry of the 343 list-4 offset slots o = (dx+3)*49 + (dy+3)*7 +
   dz+3 (dx = source - target): polar angle (index into the distinct
   angles, 49 over list 4), azimuth chi (0 on the z axis) and length (index
   into the 15 distinct lengths), as used with phi = -pi/2 */
int wbfmm_hip_offset_geometry(double *thetas, int *nth, int *th_idx,
                              double *chi, double *rr, int *nr, int *rix)
{
  int dx, dy, dz, k;
  *nth = 0;
  *nr = 0;
  for (dx = -3; dx <= 3; dx++)
    for (dy = -3; dy <= 3; dy++)
      for (dz = -3; dz <= 3; dz++) {
        int o = (dx + 3) * 49 + (dy + 3) * 7 + dz + 3;
        double th, ph, ch, r;
        th_idx[o] = rix[o] = 0;
        chi[o] = 0.0;
        if (abs(dx) <= 1 && abs(dy) <= 1 && abs(dz) <= 1) continue;
        r = sqrt((double)(dx * dx + dy * dy + dz * dz));
        dir_angles(-dx, -dy, -dz, &th, &ph, &ch);
        chi[o] = (dx == 0 && dy == 0) ? 0.0 : ch;
        for (k = 0; k < *nth; k++)
          if (fabs(thetas[k] - th) < 1e-12) break;
        if (k == *nth) thetas[(*nth)++] = th;
        th_idx[o] = k;
        for (k = 0; k < *nr; k++)
          if (rr[k] == r) break;
        if (k == *nr) rr[(*nr)++] = r;
        rix[o] = k;
      }
  return 0;
}

typedef struct {
  double *host;
  size_t nr;
  int L;
  const double *theta;
} hjob;

static void h_one(int i, void *c)
{
  hjob *j = (hjob *)c;
  H_rotation(j->host + (size_t)i * j->nr, j->L, j->theta[i]);
}

int wbfmm_hip_H_tables(int L, int nth, const double *theta, double **dev,
                       size_t *stride)
{
  size_t nr = rot_size(L);
  *stride = nr;
  double *host;
  int i, rc;
  if (L < 0 || L > 64 || nth < 1) return WBFMM_HIP_ERR_ARG;
  host = malloc(sizeof(double) * nr * nth);
  if (!host) return WBFMM_HIP_ERR_NOMEM;
  {
    hjob hj = {host, nr, L, theta};
    par_for(nth, h_one, &hj);
  }
  (void)i;
  rc = wbfmm_hip_malloc((void **)dev, sizeof(double) * nr * nth);
  if (!rc) rc = wbfmm_hip_memcpy(*dev, host, sizeof(double) * nr * nth);
  free(host);
  return rc;
}

/* index of (E|F)_{l n}^m, l >= n >= m >= 0 (src/wbfmm-private.h:149-157) */
#define CXI(l, m, n) ((size_t)(l) * ((l) + 1) * ((l) + 2) / 6 + (size_t)(n) * ((n) + 1) / 2 + (m))

/* The recursion loses about log10(amplification) digits to cancellation,
   which grows with |kr|: in x87 extended precision (64-bit mantissa) the
   tables stay within 2e-13 of the quad-precision ones for |kr| <= 20 at
   orders up to 37, at a twentieth of the cost; beyond that quad precision
   is used. */
#define CR_T long double
#define CR_NAME(x) x##_ld
#define CR_C(x) x##L
#define CR_SQRT sqrtl
#define CR_FABS fabsl
#define CR_SIN sinl
#define CR_COS cosl
#include "hel_coax.inc"
#undef CR_T
#undef CR_NAME
#undef CR_C
#undef CR_SQRT
#undef CR_FABS
#undef CR_SIN
#undef CR_COS

#define CR_T __float128
#define CR_NAME(x) x##_q
#define CR_C(x) x##Q
#define CR_SQRT sqrtq
#define CR_FABS fabsq
#define CR_SIN sinq
#define CR_COS cosq
#include "hel_coax.inc"
#undef CR_T
#undef CR_NAME
#undef CR_C
#undef CR_SQRT
#undef CR_FABS
#undef CR_SIN
#undef CR_COS

int wbfmm_hip_hel_coaxial_table(int kind, int L, double kr, double *E)
{
  if (fabs(kr) <= 20.0) return coaxial_ld(kind, L, kr, E);
  return coaxial_q(kind, L, kr, E);
}

typedef struct {
  int kind, L, rc;
  const double *kr;
  double *E;
  size_t stride;
} cxjob;

static void cx_one(int i, void *c)
{
  cxjob *j = (cxjob *)c;
  int rc = wbfmm_hip_hel_coaxial_table(j->kind, j->L, j->kr[i], j->E + (size_t)i * j->stride);
  if (rc) j->rc = rc;
}

int wbfmm_hip_hel_coaxial_tables(int kind, int L, int n, const double *kr, double *E)
{
  cxjob j = {kind, L, 0, kr, E, 2 * CXI(L + 1, 0, 0)};
  par_for(n, cx_one, &j);
  return j.rc;
}

/* ------------------------------------------------------------------
 * MFMA form of the Laplace M2M and L2L (src/shift-laplace.c:225-789):
 * the same split and bins as the M2L plan; the 8 child directions use the
 * polar angles acos(1/sqrt 3) (index 0) and acos(-1/sqrt 3) (index 1).
 *   A1, A3: 2 x KT x 64 tiles of the two angles;
 *   CSup, CSdn: 8 x (N+1) x 2, cos and sin of m chi of the child -> parent
 *     (M2M) and parent -> child (L2L) directions of child c.
 * ------------------------------------------------------------------ */
int wbfmm_hip_lap_shift_tables(const wbfmm_m2l_plan_t *P, double *A1, double *A3,
                               double *CSup, double *CSdn)
{
  packing p13;
  hcache hc;
  double *T, th[2];
  int N = P->N, S = 16 * P->NB, c, m, k;
  size_t per = (size_t)P->KT * 64;
  do_pack(N, 0, &p13);
  T = calloc((size_t)S * S, sizeof(double));
  if (!T) return WBFMM_HIP_ERR_NOMEM;
  th[0] = acos(1.0 / sqrt(3.0));
  th[1] = acos(-1.0 / sqrt(3.0));
  hc.n = 0;
  hc.L = N;
  for (k = 0; k < 2; k++)
    theta_tiles(P, &p13, cached_H(&hc, th[k]), A1 + k * per, A3 + k * per, T);
  for (k = 0; k < hc.n; k++) free(hc.H[k]);
  free(T);
  for (c = 0; c < 8; c++) {
    double ex = 0.5 - (c & 1), ey = 0.5 - ((c >> 1) & 1);
    double cu = atan2(ey, ex), cd = atan2(-ey, -ex);
    for (m = 0; m <= N; m++) {
      CSup[((size_t)c * (N + 1) + m) * 2] = cos(m * cu);
      CSup[((size_t)c * (N + 1) + m) * 2 + 1] = sin(m * cu);
      CSdn[((size_t)c * (N + 1) + m) * 2] = cos(m * cd);
      CSdn[((size_t)c * (N + 1) + m) * 2 + 1] = sin(m * cd);
    }
  }
  return 0;
}

/* stage 2 of the M2M (kind 0: S|S, c (-t)^(n-nu), nu <= n) or L2L (kind 1:
   R|R, c t^(nu-n), nu >= n) at distance t = sqrt(3)/2 w_child, same block
   for both parts (src/laplace.c:542-640) */
/* a^n b^nu scaling of the shift entries (1, 1 for the reference's
   coefficients) */
static int shift_coax(const wbfmm_m2l_plan_t *P, int kind, double t, double a, double b,
                      double *A2);

int wbfmm_hip_lap_shift_coax(const wbfmm_m2l_plan_t *P, int kind, double t, double *A2)
{
  return shift_coax(P, kind, t, 1.0, 1.0, A2);
}

/* Coefficients scaled by the box width (M_n / w^n, L_n w^(n+1)): with
   t = sqrt(3)/2 w_child and w_parent = 2 w_child the shifts no longer
   depend on the level,
     M2M  c0 (-sqrt3/2)^(n-nu) 2^-n       (M~_parent from M~_child)
     L2L  c1 (sqrt3/2)^(nu-n) 2^-(nu+1)   (L~_child from L~_parent) */
int wbfmm_hip_lap_shift_coax_unit(const wbfmm_m2l_plan_t *P, int kind, double *A2)
{
  const double t = 0.5 * sqrt(3.0);
  if (kind == 0) return shift_coax(P, 0, t, 0.5, 1.0, A2);
  if (kind == 1) return shift_coax(P, 1, t, 1.0, 0.5, A2);
  return WBFMM_HIP_ERR_ARG;
}

static int shift_coax(const wbfmm_m2l_plan_t *P, int kind, double t, double a, double b,
                      double *A2)
{
  packing p2;
  int N = P->N, S = 16 * P->NB, m, n, nu, part;
  double *T;
  if (kind != 0 && kind != 1) return WBFMM_HIP_ERR_ARG;
  do_pack(N, 1, &p2);
  T = calloc((size_t)S * S, sizeof(double));
  if (!T) return WBFMM_HIP_ERR_NOMEM;
  for (m = 0; m <= N; m++)
    for (part = 0; part < (m ? 2 : 1); part++)
      for (n = m; n <= N; n++)
        for (nu = m; nu <= N; nu++) {
          double v;
          if (kind == 0 && nu > n) continue;
          if (kind == 1 && nu < n) continue;
          v = kind == 0 ? lap_cfft(0, n, nu, m) * pow(-t, (double)(n - nu))
                        : lap_cfft(1, n, nu, m) * pow(t, (double)(nu - n));
          if (a != 1.0) v *= pow(a, (double)n);
          if (b != 1.0) v *= pow(b, (double)(nu + (kind == 1)));
          T[(size_t)slot2(&p2, N, n, m, part) * S + slot2(&p2, N, nu, m, part)] = v;
        }
  put_tiles(A2, T, P);
  free(T);
  return 0;
}

/* ------------------------------------------------------------------
 * Helmholtz M2L on matrix cores, with the Laplace plan of the same order.
 *
 * The reference translates a complex expansion x by rotate (H(theta) with
 * the phases e^{i m chi} in and i^nu out), coaxial S|R, rotate back (phases
 * (-i)^m in and e^{-i nu chi} out) (src/pass.c:56-105).  The i^nu of the
 * first rotation and the (-i)^m of the second meet across the coaxial
 * stage, which keeps the order m, and cancel.  In the basis
 *   P_m = x'_m + x'_{-m},  Q_m = x'_m - x'_{-m}   (m >= 0, P_0 = 2 x'_0)
 * of the phase-rotated x' the real H(theta) of degree n splits, by its
 * symmetry h(-nu,-m) = h(nu,m), into a cos block
 *   A_{nu m} = h(nu,m) + h(-nu,m)   (m > 0),   A_{nu 0} = h(nu,0)
 * on P (nu, m = 0..n) and a sin block B_{nu m} = h(nu,m) - h(-nu,m) on Q
 * (nu, m = 1..n): the blocks of the Laplace plan.  The coaxial table of
 * order m acts alike on P and Q.  Both rotations use the same tiles; the
 * output is y_nu = (P + Q)/2, y_{-nu} = (P - Q)/2, times e^{-i nu chi}.
 * ------------------------------------------------------------------ */

/* the distinct polar angles of the list-4 offsets in the plan's order */
static int plan_thetas(double *ths)
{
  int dx, dy, dz, n = 0;
  for (dx = -3; dx <= 3; dx++)
    for (dy = -3; dy <= 3; dy++)
      for (dz = -3; dz <= 3; dz++) {
        double th, ph, ch;
        int k;
        if (abs(dx) <= 1 && abs(dy) <= 1 && abs(dz) <= 1) continue;
        dir_angles(-dx, -dy, -dz, &th, &ph, &ch);
        for (k = 0; k < n; k++)
          if (fabs(ths[k] - th) < 1e-12) break;
        if (k == n) ths[n++] = th;
      }
  return n;
}

typedef struct {
  const wbfmm_m2l_plan_t *P;
  const packing *p13;
  const double *ths;
  double *A;
  int rc;
} hel_rot_job;

static void hel_rot_one(int k, void *c)
{
  hel_rot_job *j = (hel_rot_job *)c;
  const wbfmm_m2l_plan_t *P = j->P;
  int N = P->N, S = 16 * P->NB, n, nu, m;
  double *H = malloc(sizeof(double) * rot_size(N));
  double *T = calloc((size_t)S * S, sizeof(double));
  if (!H || !T) {
    j->rc = WBFMM_HIP_ERR_NOMEM;
    free(H);
    free(T);
    return;
  }
  H_rotation(H, N, j->ths[k]);
  for (n = 0; n <= N; n++) {
    for (nu = 0; nu <= n; nu++)
      for (m = 0; m <= n; m++)
        T[(size_t)slot13(j->p13, N, n, nu, 0) * S + slot13(j->p13, N, n, m, 0)] =
            m == 0 ? H[RIDX(nu, 0, n)] : H[RIDX(nu, m, n)] + H[RIDX(-nu, m, n)];
    for (nu = 1; nu <= n; nu++)
      for (m = 1; m <= n; m++)
        T[(size_t)slot13(j->p13, N, n, nu, 1) * S + slot13(j->p13, N, n, m, 1)] =
            H[RIDX(nu, m, n)] - H[RIDX(-nu, m, n)];
  }
  put_tiles(j->A + (size_t)k * P->KT * 64, T, P);
  free(H);
  free(T);
}

/* rotation tiles of the Helmholtz translations for plan P at the nth
   polar angles ths: nth x KT x 64, the same for both rotations */
int wbfmm_hip_hel_rot_tiles(const wbfmm_m2l_plan_t *P, int nth, const double *ths, double *A)
{
  packing p13;
  hel_rot_job j;
  if (nth < 1) return WBFMM_HIP_ERR_ARG;
  if (do_pack(P->N, 0, &p13) != P->NB) return WBFMM_HIP_ERR_ARG;
  j.P = P;
  j.p13 = &p13;
  j.ths = ths;
  j.A = A;
  j.rc = 0;
  par_for(nth, hel_rot_one, &j);
  return j.rc;
}

/* rotation tiles of the Helmholtz M2L for plan P: NTH x KT x 64 at the
   plan's polar angles */
int wbfmm_hip_hel_m2l_rot_tiles(const wbfmm_m2l_plan_t *P, double *A)
{
  double ths[64];
  if (plan_thetas(ths) != P->NTH) return WBFMM_HIP_ERR_ARG;
  return wbfmm_hip_hel_rot_tiles(P, P->NTH, ths, A);
}

/* coaxial tiles of the Helmholtz M2L for plan P from nr host S|R tables E
   of order L >= P->N (2 doubles per entry CXI(l, m, n), table i at
   E + i * 2 CXI(L+1, 0, 0)): A2 + (2 i + c) KT 64 holds the real (c = 0)
   or imaginary (c = 1) part, on the blocks of order m for P and Q, with
   the reference's (E|F) indexing: input degree l, output n, entry
   CXI(max(l,n), m, min(l,n)), negated when l < n and n + l is odd */
int wbfmm_hip_hel_m2l_coax_tiles(const wbfmm_m2l_plan_t *P, int L, int nr, const double *E,
                                 double *A2)
{
  return wbfmm_hip_hel_coax_tiles_scaled(P, L, nr, E, NULL, NULL, A2);
}

/* the same tiles with output degree n scaled by sout[n] and input degree l
   by sin[l] (NULL: 1), for coefficients held as M_n / s_n or L_n s_n */
int wbfmm_hip_hel_coax_tiles_scaled(const wbfmm_m2l_plan_t *P, int L, int nr, const double *E,
                                    const double *sout, const double *sin, double *A2)
{
  packing p2;
  int N = P->N, S = 16 * P->NB, i, c, m, part, n, l;
  size_t ne = CXI(L + 1, 0, 0), per = (size_t)P->KT * 64;
  double *T;
  if (L < N) return WBFMM_HIP_ERR_ORDER;
  if (do_pack(N, 1, &p2) != P->NB) return WBFMM_HIP_ERR_ARG;
  T = calloc((size_t)S * S, sizeof(double));
  if (!T) return WBFMM_HIP_ERR_NOMEM;
  for (i = 0; i < nr; i++)
    for (c = 0; c < 2; c++) {
      memset(T, 0, sizeof(double) * (size_t)S * S);
      for (m = 0; m <= N; m++)
        for (part = 0; part < (m ? 2 : 1); part++)
          for (n = m; n <= N; n++)
            for (l = m; l <= N; l++) {
              size_t idx = l >= n ? CXI(l, m, n) : CXI(n, m, l);
              double v = E[(size_t)i * 2 * ne + 2 * idx + c];
              if (l < n && ((n + l) & 1)) v = -v;
              if (sout) v *= sout[n];
              if (sin) v *= sin[l];
              T[(size_t)slot2(&p2, N, n, m, part) * S + slot2(&p2, N, l, m, part)] = v;
            }
      put_tiles(A2 + (size_t)(2 * i + c) * per, T, P);
    }
  free(T);
  return 0;
}
