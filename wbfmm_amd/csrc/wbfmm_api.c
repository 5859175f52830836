/*
 * wbfmm_api.c -- host side (C) of the drop-in library: the reference's
 * wbfmm_tree_* / upward-downward pass / target-field API (wbfmm.h),
 * implemented over the thin HIP C ABI of wbfmm_hip.h.
 *
 * The reference keeps everything on the host and runs each pass as C
 * loops over boxes (src/boxes.c, src/pass-laplace.c, src/local-laplace.c).
 * Here the public structs keep their reference layout and host-visible
 * fields, while points, box ranges and expansion coefficients live in
 * GPU memory; each API call queues the corresponding kernels on the
 * library stream.  Calls that hand results back to host memory wait for
 * them.
 */

#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <pthread.h>
#include <setjmp.h>
#include <stdlib.h>
#include <string.h>

#include "wbfmm.h"
#include "wbfmm_hip.h"

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

/* Abort on misuse, like the reference's g_error(), unless the calling
   thread runs under wbfmm_guarded_call (the serving engine's workers):
   then the locks this library holds in the thread are released and the
   guarded call returns the failure as a status instead. */
static __thread jmp_buf *t_recover = NULL;
static __thread int t_recover_rc = 0;
static __thread char t_fatal_msg[512];
#define MAX_HELD 8
static __thread pthread_mutex_t *t_held[MAX_HELD];
static __thread int t_nheld = 0;

static void mu_lock(pthread_mutex_t *m)
{
  pthread_mutex_lock(m);
  if (t_nheld < MAX_HELD) t_held[t_nheld] = m;
  t_nheld++;
}

static void mu_unlock(pthread_mutex_t *m)
{
  int i;
  for (i = (t_nheld < MAX_HELD ? t_nheld : MAX_HELD) - 1; i >= 0; i--)
    if (t_held[i] == m) {
      for (; i + 1 < t_nheld && i + 1 < MAX_HELD; i++) t_held[i] = t_held[i + 1];
      break;
    }
  t_nheld--;
  pthread_mutex_unlock(m);
}

static void fatal_exit(int rc) __attribute__((noreturn));
static void fatal_exit(int rc)
{
  if (t_recover) {
    jmp_buf *jb = t_recover;
    while (t_nheld > 0) {
      t_nheld--;
      if (t_nheld < MAX_HELD) pthread_mutex_unlock(t_held[t_nheld]);
    }
    t_recover_rc = rc ? rc : WBFMM_HIP_ERR_ARG;
    longjmp(*jb, 1);
  }
  abort();
}

#define WBFMM_FATAL_RC(rc, ...)                                                   \
  do {                                                                            \
    int _n = snprintf(t_fatal_msg, sizeof(t_fatal_msg), "wbfmm: %s: ", __func__); \
    if (_n > 0 && _n < (int)sizeof(t_fatal_msg))                                  \
      snprintf(t_fatal_msg + _n, sizeof(t_fatal_msg) - _n, __VA_ARGS__);          \
    fprintf(stderr, "%s\n", t_fatal_msg);                                         \
    fatal_exit(rc);                                                               \
  } while (0)
#define WBFMM_FATAL(...) WBFMM_FATAL_RC(WBFMM_HIP_ERR_ARG, __VA_ARGS__)

#define HIPOK(call)                                                              \
  do {                                                                           \
    int _rc = (call);                                                            \
    if (_rc != 0)                                                                \
      WBFMM_FATAL_RC(_rc, "%s failed (%d): %s", #call, _rc, wbfmm_hip_last_error()); \
  } while (0)

int wbfmm_guarded_call(int (*fn)(void *), void *arg, char *msg, size_t msglen)
{
  jmp_buf jb;
  jmp_buf *volatile prev = t_recover;
  volatile int rc = 0;
  if (setjmp(jb) == 0) {
    t_recover = &jb;
    rc = fn(arg);
  } else {
    rc = t_recover_rc;
    if (msg && msglen) snprintf(msg, msglen, "%s", t_fatal_msg);
  }
  t_recover = prev;
  return rc;
}

/* ------------------------------------------------------------------ */
/* internal state behind the public structs                           */
/* ------------------------------------------------------------------ */

typedef struct {
  wbfmm_tree_t t; /* public part, must stay first */
  int host_mirror;
  double *d_xyz;      /* sorted source positions, 4 doubles each */
  double *d_xyz_raw;  /* unsorted staging */
  double *d_nrm;      /* sorted normals, 4 doubles each, or NULL */
  uint64_t *d_codes;  /* sorted Morton codes */
  uint32_t *d_idx;    /* sorted position -> caller index */
  uint32_t *d_start[WBFMM_TREE_MAX_DEPTH + 1], *d_end[WBFMM_TREE_MAX_DEPTH + 1];
  uint32_t *d_occ[WBFMM_TREE_MAX_DEPTH + 1];
  unsigned nocc[WBFMM_TREE_MAX_DEPTH + 1];
  /* boxes holding sources or registered targets; when no target list is
     registered every box is a possible target and these stay NULL */
  uint32_t *d_flag[WBFMM_TREE_MAX_DEPTH + 1];
  uint32_t *d_act[WBFMM_TREE_MAX_DEPTH + 1];
  unsigned nact[WBFMM_TREE_MAX_DEPTH + 1];
  int have_targets; /* d_flag/d_act describe the current points */
  size_t mps_bytes[WBFMM_TREE_MAX_DEPTH + 1], mpr_bytes[WBFMM_TREE_MAX_DEPTH + 1];
  double *d_q, *d_dip; /* strengths in sorted order */
  unsigned cap_q, cap_dip;
  int prec32; /* FP32 coefficients scaled by the box width (Laplace) */
  /* the origin in double for the device passes; the public x[] holds it
     in the tree's own precision (gfloat for wbfmm_tree_new_f), as the
     reference does */
  double x0[3];
  /* Helmholtz with prec32 (mixed precision): the wavenumber the levels
     were set up for, the levels kept in FP64 (large k a_l, or orders the
     FP32 passes do not serve), and the scale s_n of every FP32 level (host
     and device copies, orders 0..N+1): FP32 levels hold M_n / s_n and
     L_n s_n */
  double hk;
  int hk_set;
  int h64[WBFMM_TREE_MAX_DEPTH + 1];
  double *hsc[WBFMM_TREE_MAX_DEPTH + 1], *d_hsc[WBFMM_TREE_MAX_DEPTH + 1];
} tree_dev;

typedef struct {
  wbfmm_target_list_t l; /* public part, must stay first */
  double hel_k;          /* wavenumber (Helmholtz lists) */
  double *d_xt;          /* target positions (sorted), 4 doubles each */
  uint32_t *d_tid;       /* sorted -> caller index */
  uint64_t *d_codes;
} tlist_dev;

#define M2L_A2_CACHE 64

typedef struct {
  wbfmm_shift_operators_t op; /* public part, must stay first */
  double *d_rot;
  wbfmm_rot_layout_t lay;
  double *d_lap_coax;
  /* MFMA M2L plans, one per order (built on first use) */
  int m2l_have[WBFMM_HIP_MAX_ORDER + 1];
  wbfmm_m2l_plan_t *m2l_plan[WBFMM_HIP_MAX_ORDER + 1];
  void *m2l_plan_dev[WBFMM_HIP_MAX_ORDER + 1];
  double *m2l_A1[WBFMM_HIP_MAX_ORDER + 1], *m2l_A3[WBFMM_HIP_MAX_ORDER + 1],
      *m2l_CS[WBFMM_HIP_MAX_ORDER + 1];
  /* coaxial tables of the MFMA M2L (kind 2, S|R at the 15 lengths), M2M
     (kind 0, S|S) and L2L (kind 1, R|R), one per (order, kind, width) */
  int m2l_nA2;
  int m2l_A2_N[M2L_A2_CACHE], m2l_A2_kind[M2L_A2_CACHE];
  double m2l_A2_w[M2L_A2_CACHE];
  double *m2l_A2[M2L_A2_CACHE];
  /* MFMA M2M / L2L rotation tiles and child-direction phases per order */
  double *sh_A1[WBFMM_HIP_MAX_ORDER + 1], *sh_A3[WBFMM_HIP_MAX_ORDER + 1];
  double *sh_CSup[WBFMM_HIP_MAX_ORDER + 1], *sh_CSdn[WBFMM_HIP_MAX_ORDER + 1];
  /* FP32 tables per order (coefficients scaled by the box width): M2L
     A1, A3, CS and unit-width A2; M2M / L2L A1, A3, CSup, CSdn and the
     level-independent A2 of each direction */
  float *f_A1[WBFMM_HIP_MAX_ORDER + 1], *f_A3[WBFMM_HIP_MAX_ORDER + 1],
      *f_CS[WBFMM_HIP_MAX_ORDER + 1], *f_A2[WBFMM_HIP_MAX_ORDER + 1];
  float *f_shA1[WBFMM_HIP_MAX_ORDER + 1], *f_shA3[WBFMM_HIP_MAX_ORDER + 1],
      *f_CSup[WBFMM_HIP_MAX_ORDER + 1], *f_CSdn[WBFMM_HIP_MAX_ORDER + 1],
      *f_A2up[WBFMM_HIP_MAX_ORDER + 1], *f_A2dn[WBFMM_HIP_MAX_ORDER + 1];
  /* Helmholtz: wavenumber of the coaxial tables, their orders per level
     (op.L for S|R, ss_L for R|R), H tables of the 49 list-4 polar angles
     and of the two child directions (order op.Lmax), offset geometry */
  double hel_k;
  int hel_k_set;
  guint ss_L[WBFMM_TREE_MAX_DEPTH + 1];
  double *d_hel_Hth, *d_hel_Hud;
  size_t hel_nrot;
  int *d_th_idx, *d_rix;
  double *d_chi;
  /* Helmholtz M2L on matrix cores: rotation tiles per order (shared by
     the process), the host S|R tables of each level (kept for the tiles)
     with the offset lengths they follow, and the coaxial tiles of each
     level with the order they were made for */
  double *hel_A[WBFMM_HIP_MAX_ORDER + 1];
  double *sr_host[WBFMM_TREE_MAX_DEPTH + 1];
  double sr_rr[16];
  int sr_nr;
  double *hel_A2[WBFMM_TREE_MAX_DEPTH + 1];
  int hel_A2_N[WBFMM_TREE_MAX_DEPTH + 1];
  /* Helmholtz M2M / L2L on matrix cores: rotation tiles of the two child
     polar angles per order (shared by the process), the host R|R table of
     each level and its real coaxial tiles with the order they were made
     for */
  double *hel_shA[WBFMM_HIP_MAX_ORDER + 1];
  double *ss_host[WBFMM_TREE_MAX_DEPTH + 1];
  double *hel_SS2[WBFMM_TREE_MAX_DEPTH + 1];
  int hel_SS2_N[WBFMM_TREE_MAX_DEPTH + 1];
  /* mixed-precision Helmholtz: FP32 rotation tiles and phases per order,
     and per level the scaled FP32 coaxial tiles of the M2L and of the
     M2M into / L2L from the level above, with the order and wavenumber
     they were made for */
  float *f_hA[WBFMM_HIP_MAX_ORDER + 1], *f_hshA[WBFMM_HIP_MAX_ORDER + 1];
  float *f_hCS[WBFMM_HIP_MAX_ORDER + 1], *f_hCSup[WBFMM_HIP_MAX_ORDER + 1],
      *f_hCSdn[WBFMM_HIP_MAX_ORDER + 1];
  float *f_hA2[WBFMM_TREE_MAX_DEPTH + 1], *f_hUp[WBFMM_TREE_MAX_DEPTH + 1],
      *f_hDn[WBFMM_TREE_MAX_DEPTH + 1];
  int f_hA2_N[WBFMM_TREE_MAX_DEPTH + 1], f_hUp_N[WBFMM_TREE_MAX_DEPTH + 1],
      f_hDn_N[WBFMM_TREE_MAX_DEPTH + 1];
  double f_hA2_k[WBFMM_TREE_MAX_DEPTH + 1], f_hUp_k[WBFMM_TREE_MAX_DEPTH + 1],
      f_hDn_k[WBFMM_TREE_MAX_DEPTH + 1];
  double f_hD; /* box width of the domain the scaled tiles follow */
} ops_dev;

/* orders served by the MFMA M2L kernel (wbfmm_hip_lap_m2l_mfma) */
#define M2L_MFMA_MIN 2
#define M2L_MFMA_MAX 24
/* orders served by the Helmholtz matrix-core M2L, M2M and L2L */
#define HEL_MFMA_MAX 27
/* orders served by the FP32 passes (the FP64 matrix-core M2M / L2L serve
   M2L_MFMA_MIN..M2L_MFMA_MAX) */
#define SHIFT_MFMA_MAX 15
/* deepest level served by the matrix-core M2L (30-bit Morton arithmetic
   on source boxes; a dense level array below it exceeds device memory) */
#define M2L_MFMA_MAX_LEVEL 10
/* source components of a Laplace tree (nq) */
#define LAP_MAX_NQ 8

static tree_dev *TD(wbfmm_tree_t *t) { return (tree_dev *)t; }
static int level_f32(tree_dev *td, guint level);
static void hel_prepare(wbfmm_tree_t *t, double k);
/* origin of the domain cube in double, whatever the tree's precision */
#define ORIGIN(_t) (TD(_t)->x0)

static int g_lap_coax_N = -1; /* wbfmm_laplace_coaxial_translate_init */

/* ------------------------------------------------------------------ */
/* indexing (precision independent), src/indexing.c:165-427           */
/* ------------------------------------------------------------------ */

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

guint64 wbfmm_box_index(guint32 i, guint32 j, guint32 k)
{
  return spread3(i) | (spread3(j) << 1) | (spread3(k) << 2);
}

gint wbfmm_box_location(guint64 idx, guint32 *i, guint32 *j, guint32 *k)
{
  *i = compact3(idx);
  *j = compact3(idx >> 1);
  *k = compact3(idx >> 2);
  return 0;
}

gint wbfmm_point_from_index(guint64 i, guint32 *x, guint32 *y, guint32 *z)
{
  return wbfmm_box_location(i, x, y, z);
}

guint64 wbfmm_point_locate_box(guint64 x, guint level)
{
  return x >> 3 * (20 - level);
}

guint64 wbfmm_box_parent(guint64 idx) { return idx >> 3; }
guint64 wbfmm_box_first_child(guint64 idx) { return idx << 3; }

gint wbfmm_box_neighbours(guint level, guint64 idx, guint64 *neighbours)
{
  guint32 i, j, k;
  int nbox = 1 << level, n = 0, ii, jj, kk;
  wbfmm_box_location(idx, &i, &j, &k);
  for (ii = (int)i - 1; ii <= (int)i + 1; ii++) {
    if (ii < 0 || ii >= nbox) continue;
    for (jj = (int)j - 1; jj <= (int)j + 1; jj++) {
      if (jj < 0 || jj >= nbox) continue;
      for (kk = (int)k - 1; kk <= (int)k + 1; kk++) {
        if (kk < 0 || kk >= nbox) continue;
        neighbours[n++] = wbfmm_box_index(ii, jj, kk);
      }
    }
  }
  return n;
}

/* list-4 entries are (box, offset slot) pairs, as in the reference; with
   sort TRUE the entries are grouped by polar angle of the offset, the
   grouping the reference uses for its rotation tables */
typedef struct { guint64 b, slot; double key; } l4ent;

static int cmp_l4(const void *a, const void *b)
{
  const l4ent *x = a, *y = b;
  if (x->key < y->key) return -1;
  if (x->key > y->key) return 1;
  return (x->slot < y->slot) ? -1 : (x->slot > y->slot);
}

gint wbfmm_box_interaction_list_4(guint level, guint64 idx, guint64 *list,
                                  gboolean sort)
{
  guint32 i, j, k;
  int nbox = 1 << level, n = 0, ii, jj, kk, i0, j0, k0;
  l4ent ent[216];
  wbfmm_box_location(idx, &i, &j, &k);
  i0 = (int)i - (int)(i % 2) - 2;
  j0 = (int)j - (int)(j % 2) - 2;
  k0 = (int)k - (int)(k % 2) - 2;
  for (ii = i0; ii <= i0 + 5; ii++) {
    if (ii < 0 || ii >= nbox) continue;
    for (jj = j0; jj <= j0 + 5; jj++) {
      if (jj < 0 || jj >= nbox) continue;
      for (kk = k0; kk <= k0 + 5; kk++) {
        int dx = ii - (int)i, dy = jj - (int)j, dz = kk - (int)k;
        if (kk < 0 || kk >= nbox) continue;
        if (abs(dx) <= 1 && abs(dy) <= 1 && abs(dz) <= 1) continue;
        ent[n].b = wbfmm_box_index(ii, jj, kk);
        ent[n].slot = (guint64)((dx + 3) * 49 + (dy + 3) * 7 + dz + 3);
        ent[n].key = acos(-dz / sqrt((double)(dx * dx + dy * dy + dz * dz)));
        n++;
      }
    }
  }
  if (sort) qsort(ent, n, sizeof(l4ent), cmp_l4);
  for (ii = 0; ii < n; ii++) {
    list[2 * ii] = ent[ii].b;
    list[2 * ii + 1] = ent[ii].slot;
  }
  return n;
}

/* sorted distinct cos(theta) of the list-4 offsets, built once */
static double g_l4_vals[64];
static int g_l4_nv = 0;
static pthread_once_t g_l4_once = PTHREAD_ONCE_INIT;

static void l4_values_init(void)
{
  int dx, dy, dz, a, nv = 0;
  for (dx = -3; dx <= 3; dx++)
    for (dy = -3; dy <= 3; dy++)
      for (dz = -3; dz <= 3; dz++) {
        double v;
        int b, dup = 0;
        if (abs(dx) <= 1 && abs(dy) <= 1 && abs(dz) <= 1) continue;
        v = -dz / sqrt((double)(dx * dx + dy * dy + dz * dz));
        for (b = 0; b < nv; b++)
          if (fabs(g_l4_vals[b] - v) < 1e-12) dup = 1;
        if (!dup) g_l4_vals[nv++] = v;
      }
  for (a = 0; a < nv; a++) {
    int b;
    for (b = a + 1; b < nv; b++)
      if (g_l4_vals[b] > g_l4_vals[a]) {
        double tt = g_l4_vals[a];
        g_l4_vals[a] = g_l4_vals[b];
        g_l4_vals[b] = tt;
      }
  }
  g_l4_nv = nv;
}

gint wbfmm_box_interaction_index(gint i, gint j, gint k)
{
  /* index of the polar-angle class of offset (i,j,k): the reference
     returns the row of its 49-entry theta table (src/indexing.c:385-392);
     this returns the same classification as the position of cos(theta)
     among the sorted distinct values over all list-4 offsets */
  const double *vals = g_l4_vals;
  double c, r = sqrt((double)(i * i + j * j + k * k));
  int a, nv;
  pthread_once(&g_l4_once, l4_values_init);
  nv = g_l4_nv;
  if (i < -3 || i > 3 || j < -3 || j > 3 || k < -3 || k > 3)
    WBFMM_FATAL("offset (%d,%d,%d) outside the 7x7x7 grid", i, j, k);
  c = -k / r;
  for (a = 0; a < nv; a++)
    if (fabs(vals[a] - c) < 1e-12) return a;
  return -1;
}

/* ------------------------------------------------------------------ */
/* geometry helpers, src/util.c:280-319, src/boxes.c:324-339, 585-634 */
/* ------------------------------------------------------------------ */

guint64 wbfmm_point_index_3d(gdouble *x, gdouble *c, gdouble D)
{
  guint32 xi = (guint32)((x[0] - c[0]) / D * WBFMM_INDEX_SHIFT);
  guint32 yi = (guint32)((x[1] - c[1]) / D * WBFMM_INDEX_SHIFT);
  guint32 zi = (guint32)((x[2] - c[2]) / D * WBFMM_INDEX_SHIFT);
  return wbfmm_box_index(xi, yi, zi);
}

guint64 wbfmm_point_box(wbfmm_tree_t *t, guint level, gdouble *x)
{
  guint64 i = wbfmm_point_index_3d(x, ORIGIN(t), t->D);
  return wbfmm_point_locate_box(i, level);
}

gint wbfmm_box_location_from_index(guint64 idx, guint32 level, gdouble *x0,
                                   gdouble D, gdouble *x, gdouble *wb)
{
  guint32 i, j, k;
  if (!(idx < ((guint64)1 << (3 * level))))
    WBFMM_FATAL("box %llu is not at level %u", (unsigned long long)idx, level);
  *wb = D / (1 << level);
  wbfmm_box_location(idx, &i, &j, &k);
  x[0] = x0[0] + (*wb) * i;
  x[1] = x0[1] + (*wb) * j;
  x[2] = x0[2] + (*wb) * k;
  return 0;
}

gint wbfmm_tree_box_centre(wbfmm_tree_t *t, guint level, guint64 b,
                           gdouble *xb, gdouble *wb)
{
  wbfmm_box_location_from_index(b, level, ORIGIN(t),
                                t->D, xb, wb);
  xb[0] += 0.5 * (*wb);
  xb[1] += 0.5 * (*wb);
  xb[2] += 0.5 * (*wb);
  return 0;
}

gint wbfmm_points_origin_width(gdouble *x, gint str, gint n, gdouble *xmin,
                               gdouble *xmax, gdouble *D,
                               gboolean init_limits)
{
  gint i;
  if (init_limits) {
    xmin[0] = xmax[0] = x[0];
    xmin[1] = xmax[1] = x[1];
    xmin[2] = xmax[2] = x[2];
  }
  for (i = 0; i < n; i++) {
    int a;
    for (a = 0; a < 3; a++) {
      double v = x[i * str + a];
      if (v < xmin[a]) xmin[a] = v;
      if (v > xmax[a]) xmax[a] = v;
    }
  }
  *D = xmax[0] - xmin[0];
  if (xmax[1] - xmin[1] > *D) *D = xmax[1] - xmin[1];
  if (xmax[2] - xmin[2] > *D) *D = xmax[2] - xmin[2];
  return 0;
}

gint wbfmm_truncation_number(wbfmm_tree_t *t, gdouble k, guint level,
                             gdouble tol)
{
  /* Gumerov & Duraiswami, JASA 2009 (src/util.c:474-503):
     p = (p_lo^4 + p_hi^4)^(1/4) */
  const double del = 2.0;
  double ka = 0.5 * sqrt(3.0) * k * t->D / (1 << level);
  double plo = 1.0 - log(tol * (1.0 - 1.0 / del) * sqrt(1.0 - 1.0 / del)) / log(del);
  double phi = cbrt(3.0 * log(1.0 / tol));
  phi = ka + 0.5 * phi * phi * cbrt(ka);
  return (gint)ceil(sqrt(sqrt(plo * plo * plo * plo + phi * phi * phi * phi)));
}

/* ------------------------------------------------------------------ */
/* tree                                                               */
/* ------------------------------------------------------------------ */

wbfmm_tree_t *wbfmm_tree_new(gdouble *x, gdouble D, guint maxpoints)
{
  tree_dev *td = calloc(1, sizeof(tree_dev));
  wbfmm_tree_t *t = &td->t;
  double *xt = (double *)wbfmm_tree_origin(t);
  HIPOK(wbfmm_hip_init());
  td->x0[0] = x[0];
  td->x0[1] = x[1];
  td->x0[2] = x[2];
  t->problem = 0;
  t->sorted = FALSE;
  t->maxpoints = maxpoints;
  t->nq = 0;
  t->ip = malloc(sizeof(guint) * (maxpoints ? maxpoints : 1));
  t->npoints = 0;
  t->size = sizeof(double);
  t->boxes[0] = calloc(1, sizeof(wbfmm_box_t));
  t->depth = 0;
  xt[0] = x[0];
  xt[1] = x[1];
  xt[2] = x[2];
  t->D = D;
  td->host_mirror = 1;
  HIPOK(wbfmm_hip_malloc((void **)&td->d_xyz, sizeof(double) * 4 * (maxpoints ? maxpoints : 1)));
  HIPOK(wbfmm_hip_malloc((void **)&td->d_xyz_raw, sizeof(double) * 4 * (maxpoints ? maxpoints : 1)));
  HIPOK(wbfmm_hip_malloc((void **)&td->d_codes, sizeof(uint64_t) * (maxpoints ? maxpoints : 1)));
  HIPOK(wbfmm_hip_malloc((void **)&td->d_idx, sizeof(uint32_t) * (maxpoints ? maxpoints : 1)));
  return t;
}

gint wbfmm_tree_host_mirror(wbfmm_tree_t *t, gboolean on)
{
  TD(t)->host_mirror = on ? 1 : 0;
  return 0;
}

static void level_ranges(tree_dev *td, unsigned l)
{
  size_t nb = (size_t)1 << (3 * l);
  if (!td->d_start[l]) {
    HIPOK(wbfmm_hip_malloc((void **)&td->d_start[l], sizeof(uint32_t) * nb));
    HIPOK(wbfmm_hip_malloc((void **)&td->d_end[l], sizeof(uint32_t) * nb));
    HIPOK(wbfmm_hip_malloc((void **)&td->d_occ[l], sizeof(uint32_t) * nb));
  }
  HIPOK(wbfmm_hip_box_ranges(td->d_codes, td->t.npoints, l, td->d_start[l],
                             td->d_end[l]));
  HIPOK(wbfmm_hip_compact_boxes(td->d_start[l], td->d_end[l], NULL, l,
                                td->d_occ[l], &td->nocc[l]));
}

static void mirror_level(tree_dev *td, unsigned l)
{
  size_t nb = (size_t)1 << (3 * l), b;
  uint32_t *s, *e;
  wbfmm_box_t *boxes = td->t.boxes[l];
  if (!td->host_mirror || !boxes) return;
  s = malloc(sizeof(uint32_t) * nb);
  e = malloc(sizeof(uint32_t) * nb);
  HIPOK(wbfmm_hip_memcpy(s, td->d_start[l], sizeof(uint32_t) * nb));
  HIPOK(wbfmm_hip_memcpy(e, td->d_end[l], sizeof(uint32_t) * nb));
  for (b = 0; b < nb; b++) {
    boxes[b].n = e[b] - s[b];
    boxes[b].i = boxes[b].n ? s[b] : 0;
  }
  free(s);
  free(e);
}

static gint add_points_dbl(wbfmm_tree_t *t, gpointer pts, gsize pstr, gpointer normals,
                           gsize nstr, guint npts, gboolean sorted);

gint wbfmm_tree_add_points(wbfmm_tree_t *t, gpointer pts, gsize pstr,
                           gpointer normals, gsize nstr, guint npts,
                           gboolean sorted)
{
  if (t->size != sizeof(double))
    WBFMM_FATAL("mixed precision not implemented");
  return add_points_dbl(t, pts, pstr, normals, nstr, npts, sorted);
}

/* double positions (host or device, byte strides) into the device tree */
static gint add_points_dbl(wbfmm_tree_t *t, gpointer pts, gsize pstr, gpointer normals,
                           gsize nstr, guint npts, gboolean sorted)
{
  tree_dev *td = TD(t);
  unsigned nbad = 0;
  double *x0 = ORIGIN(t);
  if (npts > t->maxpoints)
    WBFMM_FATAL("too many points (%u) for tree (%u)", npts, t->maxpoints);
  /* a new point set starts a new tree: levels are rebuilt by refine and
     any registered target list must be added again */
  t->depth = 0;
  td->have_targets = 0;
  t->npoints = npts;
  t->points = (char *)pts;
  t->pstr = pstr;
  t->normals = (char *)normals;
  t->nstr = nstr;
  t->sorted = sorted;
  HIPOK(wbfmm_hip_morton(pts, pstr, npts, x0, t->D, td->d_xyz_raw,
                         td->d_codes, td->d_idx, &nbad));
  if (nbad)
    WBFMM_FATAL("%u points do not lie in box of width %g at (%g,%g,%g)", nbad,
                t->D, x0[0], x0[1], x0[2]);
  HIPOK(wbfmm_hip_sort(td->d_codes, td->d_idx, npts));
  HIPOK(wbfmm_hip_gather4(td->d_xyz_raw, td->d_idx, npts, td->d_xyz));
  if (normals) {
    if (!td->d_nrm)
      HIPOK(wbfmm_hip_malloc((void **)&td->d_nrm, sizeof(double) * 4 * (t->maxpoints ? t->maxpoints : 1)));
    HIPOK(wbfmm_hip_stage4(normals, nstr, npts, td->d_xyz_raw));
    HIPOK(wbfmm_hip_gather4(td->d_xyz_raw, td->d_idx, npts, td->d_nrm));
  }
  if (td->host_mirror && t->ip && npts)
    HIPOK(wbfmm_hip_memcpy(t->ip, td->d_idx, sizeof(uint32_t) * npts));
  t->boxes[0][0].i = 0;
  t->boxes[0][0].n = npts;
  level_ranges(td, 0);
  return 0;
}

/* Reorders the caller's npts records of psize bytes in place by the
   Morton index of their first three coordinates (xyz read at the start of
   each record, xstr bytes apart in the staging copy), keeping the order of
   equal indices (src/boxes.c:190-206 sorts with GLib's stable merge sort).
   Codes and the stable sort run on the GPU; the records move on the GPU
   when they are device memory and on the host otherwise. */
static gint sort_records(wbfmm_tree_t *t, const void *xyz, gsize xstr, gpointer pts,
                         gsize psize, guint npts)
{
  double *dx = NULL;
  uint64_t *dc = NULL;
  uint32_t *di = NULL;
  unsigned nbad = 0;
  if (npts == 0) return 0;
  if (psize < 3 * t->size) WBFMM_FATAL("point size %zu below three coordinates", (size_t)psize);
  HIPOK(wbfmm_hip_malloc((void **)&dx, sizeof(double) * 4 * npts));
  HIPOK(wbfmm_hip_malloc((void **)&dc, sizeof(uint64_t) * npts));
  HIPOK(wbfmm_hip_malloc((void **)&di, sizeof(uint32_t) * npts));
  HIPOK(wbfmm_hip_morton(xyz, xstr, npts, ORIGIN(t), t->D, dx, dc, di, &nbad));
  if (nbad)
    WBFMM_FATAL("%u points do not lie in box of width %g at (%g,%g,%g)", nbad, t->D,
                ORIGIN(t)[0], ORIGIN(t)[1], ORIGIN(t)[2]);
  HIPOK(wbfmm_hip_sort(dc, di, npts));
  if (wbfmm_hip_is_device(pts)) {
    void *tmp = NULL;
    HIPOK(wbfmm_hip_malloc(&tmp, psize * npts));
    HIPOK(wbfmm_hip_gather_records(pts, psize, di, npts, tmp));
    HIPOK(wbfmm_hip_memcpy(pts, tmp, psize * npts));
    HIPOK(wbfmm_hip_free(tmp));
  } else {
    uint32_t *hi = malloc(sizeof(uint32_t) * npts);
    char *tmp = malloc(psize * npts);
    guint i;
    if (!hi || !tmp) WBFMM_FATAL("out of host memory");
    HIPOK(wbfmm_hip_memcpy(hi, di, sizeof(uint32_t) * npts));
    for (i = 0; i < npts; i++) memcpy(tmp + (size_t)i * psize, (char *)pts + (size_t)hi[i] * psize, psize);
    memcpy(pts, tmp, psize * npts);
    free(hi);
    free(tmp);
  }
  HIPOK(wbfmm_hip_free(dx));
  HIPOK(wbfmm_hip_free(dc));
  HIPOK(wbfmm_hip_free(di));
  HIPOK(wbfmm_hip_sync());
  return 0;
}

gint wbfmm_tree_sort_points(wbfmm_tree_t *t, gpointer pts, gsize psize, guint npts)
{
  if (t->size != sizeof(double))
    WBFMM_FATAL("mixed precision not implemented (size of tree data type (%zu) not equal "
                "to size of requested target type (%zu))", (size_t)t->size, sizeof(double));
  return sort_records(t, pts, psize, pts, psize, npts);
}

gint wbfmm_tree_refine(wbfmm_tree_t *t)
{
  tree_dev *td = TD(t);
  unsigned l = t->depth + 1;
  if (l > WBFMM_TREE_MAX_DEPTH)
    WBFMM_FATAL("tree depth would exceed %d", WBFMM_TREE_MAX_DEPTH);
  if (td->host_mirror && !t->boxes[l])
    t->boxes[l] = calloc((size_t)1 << (3 * l), sizeof(wbfmm_box_t));
  t->depth = l;
  level_ranges(td, l);
  mirror_level(td, l);
  return 0;
}

gint wbfmm_tree_coefficients_zero(wbfmm_tree_t *t, guint level)
{
  tree_dev *td = TD(t);
  if (t->mps[level]) HIPOK(wbfmm_hip_memset(t->mps[level], 0, td->mps_bytes[level]));
  if (t->mpr[level]) HIPOK(wbfmm_hip_memset(t->mpr[level], 0, td->mpr_bytes[level]));
  return 0;
}

gint wbfmm_tree_coefficients_download(wbfmm_tree_t *t, guint level,
                                      gint which, gdouble *host)
{
  tree_dev *td = TD(t);
  void *src = which ? t->mpr[level] : t->mps[level];
  size_t bytes = which ? td->mpr_bytes[level] : td->mps_bytes[level];
  const int f32 = level_f32(td, level);
  size_t nb = (size_t)1 << (3 * level), C, g, c, k, es = f32 ? sizeof(float) : sizeof(double);
  double *tmp;
  if (!src) return -1;
  C = bytes / es / nb;
  tmp = malloc(nb * C * sizeof(double));
  if (!tmp) WBFMM_FATAL("out of host memory");
  if (f32 && t->problem == WBFMM_PROBLEM_HELMHOLTZ) {
    /* M_n = s_n M~_n, L_n = L~_n / s_n */
    float *ft = malloc(bytes);
    int n;
    if (!ft) WBFMM_FATAL("out of host memory");
    if (!td->hsc[level]) WBFMM_FATAL("level %u has no scales yet (no leaf expansions)", level);
    HIPOK(wbfmm_hip_memcpy(ft, src, bytes));
    for (g = 0; g < nb; g++)
      for (c = 0; c < C; c++) {
        const size_t cc = c / 2;
        n = (int)sqrt((double)cc);
        while ((size_t)(n + 1) * (n + 1) <= cc) n++;
        while ((size_t)n * n > cc) n--;
        tmp[g * C + c] = (double)ft[g * C + c] * (which ? 1.0 / td->hsc[level][n] : td->hsc[level][n]);
      }
    free(ft);
  } else if (f32) {
    /* FP32 coefficients scaled by the box width w: M_n = w^n M~_n,
       L_n = L~_n / w^(n+1) */
    float *ft = malloc(bytes);
    double w = t->D / (double)(1u << level), sc[WBFMM_HIP_MAX_ORDER + 2];
    int n;
    if (!ft) WBFMM_FATAL("out of host memory");
    HIPOK(wbfmm_hip_memcpy(ft, src, bytes));
    for (n = 0; n <= WBFMM_HIP_MAX_ORDER + 1; n++)
      sc[n] = which ? 1.0 / pow(w, n + 1) : pow(w, n);
    for (g = 0; g < nb; g++)
      for (c = 0; c < C; c++) {
        n = (int)sqrt((double)c);
        while ((size_t)(n + 1) * (n + 1) <= c) n++;
        while ((size_t)n * n > c) n--;
        tmp[g * C + c] = (double)ft[g * C + c] * sc[n];
      }
    free(ft);
  } else {
    HIPOK(wbfmm_hip_memcpy(tmp, src, bytes));
  }
  /* device box-major -> reference groups of 8 siblings
     (src/boxes.c:552-579): [group][coefficient][child]; level 0 has a
     single box */
  if (nb < 8) memcpy(host, tmp, nb * C * sizeof(double));
  else if (t->problem == WBFMM_PROBLEM_HELMHOLTZ) {
    /* complex coefficients: (re, im) pairs interleaved by child */
    size_t Cc = C / 2;
    for (g = 0; g < nb / 8; g++)
      for (k = 0; k < 8; k++)
        for (c = 0; c < Cc; c++) {
          host[((g * Cc + c) * 8 + k) * 2] = tmp[(8 * g + k) * C + 2 * c];
          host[((g * Cc + c) * 8 + k) * 2 + 1] = tmp[(8 * g + k) * C + 2 * c + 1];
        }
  } else if (t->nq > 1) {
    /* nq components (planes on the device): element (box 8g+k, coefficient
       c, component q) at ((g C + c) 8 + k) nq + q, the reference's
       coefficient stride 8 nq */
    const size_t nq = t->nq, Cq = C / nq;
    size_t q;
    for (q = 0; q < nq; q++)
      for (g = 0; g < nb / 8; g++)
        for (k = 0; k < 8; k++)
          for (c = 0; c < Cq; c++)
            host[((g * Cq + c) * 8 + k) * nq + q] = tmp[(q * nb + 8 * g + k) * Cq + c];
  } else
    for (g = 0; g < nb / 8; g++)
      for (k = 0; k < 8; k++)
        for (c = 0; c < C; c++)
          host[(g * C + c) * 8 + k] = tmp[(8 * g + k) * C + c];
  free(tmp);
  return 0;
}

gint wbfmm_tree_set_precision(wbfmm_tree_t *t, gint bits)
{
  tree_dev *td = TD(t);
  int l;
  if (bits != 32 && bits != 64) WBFMM_FATAL("precision %d is not 32 or 64", bits);
  for (l = 0; l <= WBFMM_TREE_MAX_DEPTH; l++)
    if (t->mps[l] || t->mpr[l])
      WBFMM_FATAL("set the precision before the coefficients are initialised");
  td->prec32 = bits == 32;
  return 0;
}

static void set_box_pointers(tree_dev *td, unsigned l, int which, int C)
{
  wbfmm_tree_t *t = &td->t;
  size_t nb = (size_t)1 << (3 * l), i;
  double *c = which ? (double *)t->mpr[l] : (double *)t->mps[l];
  wbfmm_box_t *boxes = t->boxes[l];
  if (!boxes) return;
  /* device storage is box-major (the reference interleaves groups of 8
     siblings, src/laplace.c:823-849; the download restores that order) */
  for (i = 0; i < nb; i++) {
    double *p = c + i * (size_t)C;
    if (which) boxes[i].mpr = p; else boxes[i].mps = p;
  }
}

static void alloc_level(tree_dev *td, unsigned l, int which, size_t bytes)
{
  wbfmm_tree_t *t = &td->t;
  void **slot = which ? &t->mpr[l] : &t->mps[l];
  size_t *have = which ? &td->mpr_bytes[l] : &td->mps_bytes[l];
  if (*slot && *have != bytes) {
    wbfmm_hip_free(*slot);
    *slot = NULL;
  }
  if (!*slot) HIPOK(wbfmm_hip_malloc(slot, bytes));
  *have = bytes;
  HIPOK(wbfmm_hip_memset(*slot, 0, bytes));
}

gint wbfmm_tree_laplace_coefficient_init(wbfmm_tree_t *t, guint l, guint nr,
                                         guint ns)
{
  tree_dev *td = TD(t);
  size_t nb = (size_t)1 << (3 * l);
  if ((int)t->nq < 1)
    WBFMM_FATAL("tree has invalid number of components in source terms (%d)", t->nq);
  if ((int)t->nq > LAP_MAX_NQ)
    WBFMM_FATAL("at most %d source components are supported (nq = %u)", LAP_MAX_NQ, t->nq);
  if (td->prec32 && t->nq != 1)
    WBFMM_FATAL("single-precision trees take one source component (nq = %u)", t->nq);
  if (t->problem != WBFMM_PROBLEM_LAPLACE) WBFMM_FATAL("tree is not a Laplace tree");
  if (l > t->depth) WBFMM_FATAL("level %u below tree depth %u", l, t->depth);
  if (ns > WBFMM_HIP_MAX_ORDER || nr > WBFMM_HIP_MAX_ORDER)
    WBFMM_FATAL("order above %d", WBFMM_HIP_MAX_ORDER);
  if (td->prec32 && td->host_mirror)
    WBFMM_FATAL("FP32 coefficients have no host box pointers (use the download)");
  t->order_s[l] = ns;
  t->order_r[l] = nr;
  {
    /* nq components: nq planes of the box-major level array, plane q at
       q 8^l (n+1)^2 (box pointers, when mirrored, point into plane 0) */
    const size_t es = td->prec32 ? sizeof(float) : sizeof(double);
    if (ns) {
      alloc_level(td, l, 0, t->nq * nb * (ns + 1) * (ns + 1) * es);
      if (td->host_mirror) set_box_pointers(td, l, 0, (ns + 1) * (ns + 1));
    }
    if (nr) {
      alloc_level(td, l, 1, t->nq * nb * (nr + 1) * (nr + 1) * es);
      if (td->host_mirror) set_box_pointers(td, l, 1, (nr + 1) * (nr + 1));
    }
  }
  return 0;
}

/* plane q of a Laplace level array (nq components) */
static double *lap_plane(wbfmm_tree_t *t, guint l, int which, unsigned q)
{
  const unsigned N = which ? t->order_r[l] : t->order_s[l];
  double *base = which ? (double *)t->mpr[l] : (double *)t->mps[l];
  return base + (size_t)q * ((size_t)1 << (3 * l)) * (N + 1) * (N + 1);
}

/* strengths (host or device, element stride) into sorted device order */
/* strengths (host or device, element stride) into sorted device order;
   ncomp doubles per point (1 Laplace, 2 Helmholtz) */
static double *gather_strengths_n(tree_dev *td, double *src, gint sstr,
                                  int dip, int ncomp)
{
  unsigned n = td->t.npoints;
  double **buf = dip ? &td->d_dip : &td->d_q;
  unsigned *cap = dip ? &td->cap_dip : &td->cap_q;
  if (!src) return NULL;
  if (*cap < ncomp * n) {
    if (*buf) wbfmm_hip_free(*buf);
    HIPOK(wbfmm_hip_malloc((void **)buf, sizeof(double) * ncomp * (n ? n : 1)));
    *cap = ncomp * n;
  }
  HIPOK(wbfmm_hip_gather_strided(src, sstr, ncomp, td->d_idx, n, *buf));
  return *buf;
}

static double *gather_strengths(tree_dev *td, double *src, gint sstr, int dip)
{
  return gather_strengths_n(td, src, sstr, dip, 1);
}

gint wbfmm_tree_laplace_leaf_expansions(wbfmm_tree_t *t, gdouble *src,
                                        gint sstr, gdouble *dipoles,
                                        gint dstr, gboolean zero_expansions,
                                        gdouble *work)
{
  tree_dev *td = TD(t);
  unsigned d = t->depth;
  double *q, *dp;
  int wave;
  (void)work;
  if ((int)t->nq < 1)
    WBFMM_FATAL("tree has invalid number of components in source terms (%d)", t->nq);
  if (t->problem != WBFMM_PROBLEM_LAPLACE) WBFMM_FATAL("tree is not a Laplace tree");
  if (!t->mps[d]) WBFMM_FATAL("coefficients not initialised at level %u", d);
  if (zero_expansions) HIPOK(wbfmm_hip_memset(t->mps[d], 0, td->mps_bytes[d]));
  if (src == NULL && dipoles == NULL) return 0;
  if (t->normals == NULL && dipoles != NULL)
    WBFMM_FATAL("no normals in tree but dipole strengths specified (dipoles != NULL)");
  /* one wavefront per box where the occupied leaves hold 4 or more points
     on average, one thread per box otherwise */
  wave = t->npoints >= 4u * td->nocc[d];
  if (td->prec32) {
    q = gather_strengths(td, src, sstr, 0);
    dp = gather_strengths(td, dipoles, dstr, 1);
    HIPOK((wave ? wbfmm_hip_lap_p2m_wave_f32 : wbfmm_hip_lap_p2m_f32)(
        td->d_occ[d], td->nocc[d], td->d_start[d], td->d_end[d], td->d_xyz, q,
        dp ? td->d_nrm : NULL, dp, d, ORIGIN(t), t->D, t->order_s[d], (float *)t->mps[d]));
    return 0;
  }
  {
    /* component c: strengths src[i sstr + c], dipoles[i dstr + c], into
       plane c (src/laplace.c:854-989 with nq components) */
    unsigned c;
    for (c = 0; c < t->nq; c++) {
      q = gather_strengths(td, src ? src + c : NULL, sstr, 0);
      dp = gather_strengths(td, dipoles ? dipoles + c : NULL, dstr, 1);
      HIPOK((wave ? wbfmm_hip_lap_p2m_wave : wbfmm_hip_lap_p2m)(
          td->d_occ[d], td->nocc[d], td->d_start[d], td->d_end[d], td->d_xyz, q,
          dp ? td->d_nrm : NULL, dp, d, ORIGIN(t), t->D, t->order_s[d],
          lap_plane(t, d, 0, c)));
    }
  }
  return 0;
}

/* ------------------------------------------------------------------ */
/* shift operators                                                    */
/* ------------------------------------------------------------------ */

gint wbfmm_shift_angle_table_init(void) { return 0; }

gint wbfmm_shift_angles_list4(gint i, gint j, gint k, gdouble *th,
                              gdouble *ph, gdouble *ch, gdouble *rs)
{
  /* angles of the translation from a source at offset (i,j,k) from the
     target (src/shift.c:587-626) */
  double ex = -i, ey = -j, ez = -k, r = sqrt(ex * ex + ey * ey + ez * ez);
  if (r == 0.0) WBFMM_FATAL("zero offset");
  *rs = r;
  *th = acos(ez / r);
  if (i == 0 && j == 0) { *ph = 0.0; *ch = 0.0; }
  else { *ph = -0.5 * M_PI; *ch = atan2(ey, ex); }
  return 0;
}

wbfmm_shift_operators_t *wbfmm_shift_operators_new(guint L, gboolean bw,
                                                   gdouble *work)
{
  ops_dev *od = calloc(1, sizeof(ops_dev));
  (void)work;
  if (L > WBFMM_HIP_MAX_ORDER) WBFMM_FATAL("order %u above %d", L, WBFMM_HIP_MAX_ORDER);
  od->op.size = sizeof(double);
  od->op.Lmax = L;
  od->op.bw = bw;
  /* the rotation tables of the vector passes (718 directions, 2.3 MB per
     1000 entries of the order's rotation size) are built on first use:
     the matrix-core passes and the Helmholtz passes have their own */
  od->lay.L = (int)L;
  od->lay.nrot = (size_t)(L + 1) * (2 * L + 1) * (2 * L + 3) / 3;
  od->op.nerot = (guint)od->lay.nrot;
  od->op.rotations = NULL;
  return &od->op;
}

/* Tables that depend on the order only are built once per process, by the
   first operator set that needs them, and shared by every operator set
   until the process ends: the Laplace matrix-core plans and tiles (FP64
   and FP32) per order, and the rotation tables and coaxial closed forms
   of the Laplace vector passes per largest order.  They are complete on
   the device before they are shared. */
static pthread_mutex_t g_tab_mu = PTHREAD_MUTEX_INITIALIZER;
static double *g_vec_rot[WBFMM_HIP_MAX_ORDER + 1], *g_vec_coax[WBFMM_HIP_MAX_ORDER + 1];
static wbfmm_rot_layout_t g_vec_lay[WBFMM_HIP_MAX_ORDER + 1];

/* rotation tables and coaxial closed forms of the Laplace vector passes */
static void ensure_lap_coax(ops_dev *od)
{
  int L = od->lay.L;
  if (od->d_rot) return;
  mu_lock(&g_tab_mu);
  if (!g_vec_rot[L]) {
    HIPOK(wbfmm_hip_rotation_tables(L, &g_vec_rot[L], &g_vec_lay[L]));
    HIPOK(wbfmm_hip_laplace_coaxial_tables(L, &g_vec_coax[L]));
  }
  od->d_rot = g_vec_rot[L];
  od->lay = g_vec_lay[L];
  od->d_lap_coax = g_vec_coax[L];
  od->op.rotations = od->d_rot;
  mu_unlock(&g_tab_mu);
}

gint wbfmm_shift_operators_free(wbfmm_shift_operators_t *op)
{
  ops_dev *od = (ops_dev *)op;
  int n;
  if (!op) return 0;
  /* the per-order tables are shared with other operator sets: not freed */
  for (n = 0; n < od->m2l_nA2; n++) wbfmm_hip_free(od->m2l_A2[n]);
  for (n = 0; n <= WBFMM_TREE_MAX_DEPTH; n++) {
    wbfmm_hip_free(od->op.SR[n]);
    wbfmm_hip_free(od->op.SS[n]);
    wbfmm_hip_free(od->hel_A2[n]);
    free(od->sr_host[n]);
    wbfmm_hip_free(od->hel_SS2[n]);
    free(od->ss_host[n]);
    wbfmm_hip_free(od->f_hA2[n]);
    wbfmm_hip_free(od->f_hUp[n]);
    wbfmm_hip_free(od->f_hDn[n]);
  }
  for (n = 0; n <= WBFMM_HIP_MAX_ORDER; n++) {
    wbfmm_hip_free(od->f_hA[n]);
    wbfmm_hip_free(od->f_hshA[n]);
    wbfmm_hip_free(od->f_hCS[n]);
    wbfmm_hip_free(od->f_hCSup[n]);
    wbfmm_hip_free(od->f_hCSdn[n]);
  }
  /* d_hel_*, hel_A and hel_shA point to the process-wide tables: not freed */
  free(od);
  return 0;
}

/* the shared per-order tables live in this set's fields */
static ops_dev g_tab;

/* one lock per order: a table build of one order holds up no other */
static pthread_mutex_t g_ord_mu[WBFMM_HIP_MAX_ORDER + 1] = {
    [0 ... WBFMM_HIP_MAX_ORDER] = PTHREAD_MUTEX_INITIALIZER};

/* the M2L plan of order N and its phases (no tiles) */
static void build_m2l_plan(ops_dev *od, int N)
{
  wbfmm_m2l_plan_t *P;
  if (od->m2l_plan[N]) return;
  P = calloc(1, sizeof(wbfmm_m2l_plan_t));
  if (!P) WBFMM_FATAL("out of host memory");
  HIPOK(wbfmm_hip_lap_m2l_plan(N, P, &od->m2l_plan_dev[N], NULL, NULL, &od->m2l_CS[N]));
  od->m2l_plan[N] = P;
}

/* the Laplace M2L rotation tiles of order N */
static void build_lap_m2l_tiles(ops_dev *od, int N)
{
  const wbfmm_m2l_plan_t *P = od->m2l_plan[N];
  wbfmm_m2l_plan_t Q;
  size_t n13;
  double *h1, *h3;
  if (od->m2l_A1[N]) return;
  n13 = (size_t)P->NTH * P->KT * 64;
  h1 = malloc(sizeof(double) * n13);
  h3 = malloc(sizeof(double) * n13);
  if (!h1 || !h3) WBFMM_FATAL("out of host memory");
  HIPOK(wbfmm_hip_lap_m2l_tables(N, &Q, h1, h3, NULL));
  HIPOK(wbfmm_hip_malloc((void **)&od->m2l_A3[N], sizeof(double) * n13));
  HIPOK(wbfmm_hip_memcpy(od->m2l_A3[N], h3, sizeof(double) * n13));
  HIPOK(wbfmm_hip_malloc((void **)&od->m2l_A1[N], sizeof(double) * n13));
  HIPOK(wbfmm_hip_memcpy(od->m2l_A1[N], h1, sizeof(double) * n13));
  free(h1);
  free(h3);
}

static void build_shift_tables(ops_dev *od, int N);
static void build_f32_tables(ops_dev *od, int N);
static void copy_order_tables(ops_dev *od, int N);

/* the plan of order N with its phases and the M2M / L2L tiles and phases
   (what the Helmholtz matrix-core passes take from it) into od */
static void ensure_plan(ops_dev *od, int N)
{
  if (od->m2l_plan[N] && od->sh_CSup[N]) return;
  mu_lock(&g_ord_mu[N]);
  build_m2l_plan(&g_tab, N);
  build_shift_tables(&g_tab, N);
  copy_order_tables(od, N);
  mu_unlock(&g_ord_mu[N]);
}

/* every matrix-core table of order N (FP64 M2L plan and tiles, M2M / L2L
   tiles, FP32 copies) into od, from the shared set */
static void ensure_order_tables(ops_dev *od, int N)
{
  if (od->m2l_have[N]) return;
  mu_lock(&g_ord_mu[N]);
  if (!g_tab.f_A1[N]) {
    build_m2l_plan(&g_tab, N);
    build_lap_m2l_tiles(&g_tab, N);
    build_shift_tables(&g_tab, N);
    build_f32_tables(&g_tab, N);
    HIPOK(wbfmm_hip_sync()); /* conversions queued on this thread's stream */
  }
  copy_order_tables(od, N);
  od->m2l_have[N] = 1;
  mu_unlock(&g_ord_mu[N]);
}

static void copy_order_tables(ops_dev *od, int N)
{
  od->m2l_plan[N] = g_tab.m2l_plan[N];
  od->m2l_plan_dev[N] = g_tab.m2l_plan_dev[N];
  od->m2l_A1[N] = g_tab.m2l_A1[N];
  od->m2l_A3[N] = g_tab.m2l_A3[N];
  od->m2l_CS[N] = g_tab.m2l_CS[N];
  od->sh_A1[N] = g_tab.sh_A1[N];
  od->sh_A3[N] = g_tab.sh_A3[N];
  od->sh_CSup[N] = g_tab.sh_CSup[N];
  od->sh_CSdn[N] = g_tab.sh_CSdn[N];
  od->f_A1[N] = g_tab.f_A1[N];
  od->f_A3[N] = g_tab.f_A3[N];
  od->f_CS[N] = g_tab.f_CS[N];
  od->f_A2[N] = g_tab.f_A2[N];
  od->f_shA1[N] = g_tab.f_shA1[N];
  od->f_shA3[N] = g_tab.f_shA3[N];
  od->f_CSup[N] = g_tab.f_CSup[N];
  od->f_CSdn[N] = g_tab.f_CSdn[N];
  od->f_A2up[N] = g_tab.f_A2up[N];
  od->f_A2dn[N] = g_tab.f_A2dn[N];
}

static void ensure_m2l_plan(ops_dev *od, int N) { ensure_order_tables(od, N); }

/* device coaxial tiles of the MFMA passes for order N: kind 2 S|R at the
   15 list-4 lengths for box width w (M2L), kind 0 S|S and kind 1 R|R at
   distance w (M2M, L2L); a few entries are kept */
static const double *mfma_coax(ops_dev *od, int N, int kind, double w)
{
  const wbfmm_m2l_plan_t *P = od->m2l_plan[N];
  size_t n = (size_t)(kind == 2 ? P->NR : 1) * P->KT * 64;
  double *h;
  int i;
  for (i = 0; i < od->m2l_nA2; i++)
    if (od->m2l_A2_N[i] == N && od->m2l_A2_kind[i] == kind && od->m2l_A2_w[i] == w)
      return od->m2l_A2[i];
  if (od->m2l_nA2 == M2L_A2_CACHE) {   /* full: drop the oldest */
    wbfmm_hip_sync();
    wbfmm_hip_free(od->m2l_A2[0]);
    memmove(od->m2l_A2_N, od->m2l_A2_N + 1, sizeof(int) * (M2L_A2_CACHE - 1));
    memmove(od->m2l_A2_kind, od->m2l_A2_kind + 1, sizeof(int) * (M2L_A2_CACHE - 1));
    memmove(od->m2l_A2_w, od->m2l_A2_w + 1, sizeof(double) * (M2L_A2_CACHE - 1));
    memmove(od->m2l_A2, od->m2l_A2 + 1, sizeof(double *) * (M2L_A2_CACHE - 1));
    od->m2l_nA2--;
  }
  h = malloc(sizeof(double) * n);
  if (!h) WBFMM_FATAL("out of host memory");
  if (kind == 2) HIPOK(wbfmm_hip_lap_m2l_coax(P, w, h));
  else HIPOK(wbfmm_hip_lap_shift_coax(P, kind, w, h));
  i = od->m2l_nA2++;
  od->m2l_A2_N[i] = N;
  od->m2l_A2_kind[i] = kind;
  od->m2l_A2_w[i] = w;
  HIPOK(wbfmm_hip_malloc((void **)&od->m2l_A2[i], sizeof(double) * n));
  HIPOK(wbfmm_hip_memcpy(od->m2l_A2[i], h, sizeof(double) * n));
  free(h);
  return od->m2l_A2[i];
}

static const double *m2l_coax(ops_dev *od, int N, double w) { return mfma_coax(od, N, 2, w); }

static void ensure_shift_tables(ops_dev *od, int N) { ensure_order_tables(od, N); }

/* rotation tiles and phases of the MFMA M2M / L2L for order N */
static void build_shift_tables(ops_dev *od, int N)
{
  const wbfmm_m2l_plan_t *P;
  size_t n13, ncs;
  double *a1, *a3, *cu, *cd;
  if (od->sh_A1[N]) return;
  P = od->m2l_plan[N];
  n13 = (size_t)2 * P->KT * 64;
  ncs = (size_t)8 * (N + 1) * 2;
  a1 = malloc(sizeof(double) * n13);
  a3 = malloc(sizeof(double) * n13);
  cu = malloc(sizeof(double) * ncs);
  cd = malloc(sizeof(double) * ncs);
  if (!a1 || !a3 || !cu || !cd) WBFMM_FATAL("out of host memory");
  HIPOK(wbfmm_hip_lap_shift_tables(P, a1, a3, cu, cd));
  HIPOK(wbfmm_hip_malloc((void **)&od->sh_A1[N], sizeof(double) * n13));
  HIPOK(wbfmm_hip_memcpy(od->sh_A1[N], a1, sizeof(double) * n13));
  HIPOK(wbfmm_hip_malloc((void **)&od->sh_A3[N], sizeof(double) * n13));
  HIPOK(wbfmm_hip_memcpy(od->sh_A3[N], a3, sizeof(double) * n13));
  HIPOK(wbfmm_hip_malloc((void **)&od->sh_CSup[N], sizeof(double) * ncs));
  HIPOK(wbfmm_hip_memcpy(od->sh_CSup[N], cu, sizeof(double) * ncs));
  HIPOK(wbfmm_hip_malloc((void **)&od->sh_CSdn[N], sizeof(double) * ncs));
  HIPOK(wbfmm_hip_memcpy(od->sh_CSdn[N], cd, sizeof(double) * ncs));
  free(a1);
  free(a3);
  free(cu);
  free(cd);
}

/* device FP32 copy of n device doubles */
static float *dev_f32(const double *d, size_t n)
{
  float *f = NULL;
  HIPOK(wbfmm_hip_malloc((void **)&f, sizeof(float) * n));
  HIPOK(wbfmm_hip_to_f32(d, f, n));
  return f;
}

/* device FP32 copy of n host doubles */
static float *host_f32(const double *h, size_t n)
{
  float *f = NULL, *t = malloc(sizeof(float) * n);
  size_t i;
  if (!t) WBFMM_FATAL("out of host memory");
  for (i = 0; i < n; i++) t[i] = (float)h[i];
  HIPOK(wbfmm_hip_malloc((void **)&f, sizeof(float) * n));
  HIPOK(wbfmm_hip_memcpy(f, t, sizeof(float) * n));
  free(t);
  return f;
}

static void ensure_f32_tables(ops_dev *od, int N) { ensure_order_tables(od, N); }

/* FP32 tables of order N for coefficients scaled by the box width (after
   the FP64 plan and shift tables) */
static void build_f32_tables(ops_dev *od, int N)
{
  const wbfmm_m2l_plan_t *P;
  size_t n13, ncs, na2;
  double *h;
  if (od->f_A1[N]) return;
  P = od->m2l_plan[N];
  n13 = (size_t)P->NTH * P->KT * 64;
  ncs = (size_t)343 * (N + 1) * 2;
  od->f_A1[N] = dev_f32(od->m2l_A1[N], n13);
  od->f_A3[N] = dev_f32(od->m2l_A3[N], n13);
  od->f_CS[N] = dev_f32(od->m2l_CS[N], ncs);
  od->f_shA1[N] = dev_f32(od->sh_A1[N], (size_t)2 * P->KT * 64);
  od->f_shA3[N] = dev_f32(od->sh_A3[N], (size_t)2 * P->KT * 64);
  od->f_CSup[N] = dev_f32(od->sh_CSup[N], (size_t)8 * (N + 1) * 2);
  od->f_CSdn[N] = dev_f32(od->sh_CSdn[N], (size_t)8 * (N + 1) * 2);
  na2 = (size_t)P->NR * P->KT * 64;
  h = malloc(sizeof(double) * na2);
  if (!h) WBFMM_FATAL("out of host memory");
  HIPOK(wbfmm_hip_lap_m2l_coax(P, 1.0, h));
  od->f_A2[N] = host_f32(h, na2);
  HIPOK(wbfmm_hip_lap_shift_coax_unit(P, 0, h));
  od->f_A2up[N] = host_f32(h, (size_t)P->KT * 64);
  HIPOK(wbfmm_hip_lap_shift_coax_unit(P, 1, h));
  od->f_A2dn[N] = host_f32(h, (size_t)P->KT * 64);
  free(h);
}

static int mfma_order(int n) { return n >= M2L_MFMA_MIN && n <= SHIFT_MFMA_MAX; }

gint wbfmm_laplace_coaxial_translate_init(gint N)
{
  if (N > g_lap_coax_N) g_lap_coax_N = N;
  return 0;
}

/* ------------------------------------------------------------------ */
/* Laplace passes                                                     */
/* ------------------------------------------------------------------ */

/* FP64 M2M of one component from `level' (in) into level-1 (out) */
static void lap_m2m_plane(wbfmm_tree_t *t, ops_dev *od, guint level, const double *in,
                          double *out)
{
  tree_dev *td = TD(t);
  if (t->order_s[level] == t->order_s[level - 1] && (int)t->order_s[level] >= M2L_MFMA_MIN &&
      (int)t->order_s[level] <= M2L_MFMA_MAX) {
    /* matrix-core form: 16 parents per wavefront (8 wavefronts above 15) */
    int N = (int)t->order_s[level];
    double tt = 0.5 * sqrt(3.0) * t->D / (double)(1u << level);
    ensure_shift_tables(od, N);
    HIPOK(wbfmm_hip_lap_shift_mfma(0, td->d_occ[level - 1], td->nocc[level - 1],
                                   td->d_start[level], td->d_end[level], NULL, N, in, out,
                                   od->m2l_plan_dev[N], od->sh_A1[N], od->sh_A3[N],
                                   mfma_coax(od, N, 0, tt), od->sh_CSup[N], od->m2l_plan[N]));
    return;
  }
  ensure_lap_coax(od);
  HIPOK(wbfmm_hip_lap_m2m(td->d_occ[level - 1], td->nocc[level - 1],
                          td->d_start[level], td->d_end[level], level, t->D,
                          t->order_s[level], t->order_s[level - 1], in, out,
                          od->d_rot, &od->lay, od->d_lap_coax));
}

gint wbfmm_laplace_upward_pass(wbfmm_tree_t *t, wbfmm_shift_operators_t *op,
                               guint level, gdouble *work)
{
  tree_dev *td = TD(t);
  ops_dev *od = (ops_dev *)op;
  (void)work;
  if (level <= 1) WBFMM_FATAL("level %u <= 1", level);
  if (t->problem != WBFMM_PROBLEM_LAPLACE) WBFMM_FATAL("tree is not a Laplace tree");
  if (!t->mps[level] || !t->mps[level - 1])
    WBFMM_FATAL("coefficients not initialised at levels %u, %u", level, level - 1);
  if ((int)t->order_s[level] > od->lay.L || (int)t->order_s[level - 1] > od->lay.L)
    WBFMM_FATAL("operators built to order %d, tree needs %u", od->lay.L,
                t->order_s[level] > t->order_s[level - 1] ? t->order_s[level] : t->order_s[level - 1]);
  if (td->prec32) {
    int N = (int)t->order_s[level];
    if (t->order_s[level - 1] != t->order_s[level] || !mfma_order(N))
      WBFMM_FATAL("FP32 passes need equal orders in %d..%d", M2L_MFMA_MIN, SHIFT_MFMA_MAX);
    ensure_f32_tables(od, N);
    HIPOK(wbfmm_hip_lap_shift_mfma_f32(0, td->d_occ[level - 1], td->nocc[level - 1],
                                       td->d_start[level], td->d_end[level], NULL, N,
                                       (const float *)t->mps[level], (float *)t->mps[level - 1],
                                       od->m2l_plan_dev[N], od->f_shA1[N], od->f_shA3[N],
                                       od->f_A2up[N], od->f_CSup[N], od->m2l_plan[N]));
    return 0;
  }
  {
    unsigned c;
    for (c = 0; c < t->nq; c++) /* each source component's plane */
      lap_m2m_plane(t, od, level, lap_plane(t, level, 0, c), lap_plane(t, level - 1, 0, c));
  }
  return 0;
}

/* all boxes of a level, used as the target list when no target list is
   registered (every box can hold a target).  The list only grows; a
   replaced list is kept, since launches on other streams may still read
   it, so at most twice the largest list is held. */
static uint32_t *g_all_boxes = NULL;
static unsigned g_all_n = 0;
static uint32_t *g_all_old[WBFMM_TREE_MAX_DEPTH + 1];
static int g_all_nold = 0;
static pthread_mutex_t g_all_mu = PTHREAD_MUTEX_INITIALIZER;

static uint32_t *all_boxes(unsigned level, unsigned *n)
{
  unsigned nb = 1u << (3 * level);
  uint32_t *list;
  mu_lock(&g_all_mu);
  if (nb > g_all_n) {
    uint32_t *h = malloc(sizeof(uint32_t) * nb), *d = NULL;
    unsigned i;
    if (!h) WBFMM_FATAL("out of host memory");
    for (i = 0; i < nb; i++) h[i] = i;
    HIPOK(wbfmm_hip_malloc((void **)&d, sizeof(uint32_t) * nb));
    HIPOK(wbfmm_hip_memcpy(d, h, sizeof(uint32_t) * nb)); /* synchronous */
    free(h);
    if (g_all_boxes) g_all_old[g_all_nold++] = g_all_boxes;
    g_all_boxes = d;
    g_all_n = nb;
  }
  list = g_all_boxes;
  mu_unlock(&g_all_mu);
  *n = nb;
  return list;
}

/* FP64 L2L of one component from the parents tl at `level' (in) into
   their children at level+1 (out, added) */
static void lap_l2l_plane(wbfmm_tree_t *t, ops_dev *od, guint level, const uint32_t *tl,
                          unsigned nt, const double *in, double *out)
{
  tree_dev *td = TD(t);
  if (t->order_r[level] == t->order_r[level + 1] && (int)t->order_r[level] >= M2L_MFMA_MIN &&
      (int)t->order_r[level] <= M2L_MFMA_MAX) {
    int N = (int)t->order_r[level];
    double tt = 0.5 * sqrt(3.0) * t->D / (double)(1u << (level + 1));
    ensure_shift_tables(od, N);
    HIPOK(wbfmm_hip_lap_shift_mfma(1, tl, nt, NULL, NULL,
                                   td->have_targets ? td->d_flag[level + 1] : NULL, N, in, out,
                                   od->m2l_plan_dev[N], od->sh_A1[N], od->sh_A3[N],
                                   mfma_coax(od, N, 1, tt), od->sh_CSdn[N], od->m2l_plan[N]));
    return;
  }
  ensure_lap_coax(od);
  HIPOK(wbfmm_hip_lap_l2l(tl, nt, td->have_targets ? td->d_flag[level + 1] : NULL,
                          level, t->D, t->order_r[level], t->order_r[level + 1], in, out,
                          od->d_rot, &od->lay, od->d_lap_coax));
}

static gint lap_downward(wbfmm_tree_t *t, wbfmm_shift_operators_t *op,
                         guint level)
{
  tree_dev *td = TD(t);
  ops_dev *od = (ops_dev *)op;
  uint32_t *tl;
  unsigned nt;
  if (level <= 1) WBFMM_FATAL("level %u <= 1", level);
  if (t->problem != WBFMM_PROBLEM_LAPLACE) WBFMM_FATAL("tree is not a Laplace tree");
  if (!t->mpr[level] || !t->mps[level])
    WBFMM_FATAL("coefficients not initialised at level %u", level);
  if ((int)t->order_s[level] > od->lay.L || (int)t->order_r[level] > od->lay.L)
    WBFMM_FATAL("operators built to order %d", od->lay.L);
  if (td->have_targets) {
    tl = td->d_act[level];
    nt = td->nact[level];
  } else {
    tl = all_boxes(level, &nt);
  }
  if (td->prec32) {
    int N = (int)t->order_s[level];
    unsigned cstart[8], ccount[8];
    uint32_t *cls;
    if (t->order_r[level] != t->order_s[level] || !mfma_order(N) ||
        (level < t->depth && t->order_r[level + 1] != t->order_r[level]))
      WBFMM_FATAL("FP32 passes need equal orders in %d..%d", M2L_MFMA_MIN, SHIFT_MFMA_MAX);
    if (level > M2L_MFMA_MAX_LEVEL)
      WBFMM_FATAL("FP32 passes serve levels up to %d", M2L_MFMA_MAX_LEVEL);
    ensure_f32_tables(od, N);
    HIPOK(wbfmm_hip_malloc((void **)&cls, sizeof(uint32_t) * (nt ? nt : 1)));
    HIPOK(wbfmm_hip_split_classes(tl, nt, cls, cstart, ccount));
    HIPOK(wbfmm_hip_lap_m2l_mfma_f32(cls, cstart, ccount, td->d_start[level],
                                     td->d_end[level], level, N,
                                     (const float *)t->mps[level], (float *)t->mpr[level],
                                     od->m2l_plan_dev[N], od->f_A1[N], od->f_A3[N],
                                     od->f_A2[N], od->f_CS[N], od->m2l_plan[N]));
    wbfmm_hip_free(cls);
    if (level == t->depth) return 0;
    if (!t->mpr[level + 1]) WBFMM_FATAL("coefficients not initialised at level %u", level + 1);
    HIPOK(wbfmm_hip_lap_shift_mfma_f32(1, tl, nt, NULL, NULL,
                                       td->have_targets ? td->d_flag[level + 1] : NULL, N,
                                       (const float *)t->mpr[level], (float *)t->mpr[level + 1],
                                       od->m2l_plan_dev[N], od->f_shA1[N], od->f_shA3[N],
                                       od->f_A2dn[N], od->f_CSdn[N], od->m2l_plan[N]));
    return 0;
  }
  if (t->order_s[level] == t->order_r[level] && level <= M2L_MFMA_MAX_LEVEL &&
      (int)t->order_s[level] >= M2L_MFMA_MIN && (int)t->order_s[level] <= M2L_MFMA_MAX) {
    /* matrix-core form: targets grouped by Morton child digit (once for
       every component) */
    int N = (int)t->order_s[level];
    unsigned cstart[8], ccount[8], c;
    uint32_t *cls;
    ensure_m2l_plan(od, N);
    HIPOK(wbfmm_hip_malloc((void **)&cls, sizeof(uint32_t) * (nt ? nt : 1)));
    HIPOK(wbfmm_hip_split_classes(tl, nt, cls, cstart, ccount));
    for (c = 0; c < t->nq; c++)
      HIPOK(wbfmm_hip_lap_m2l_mfma(cls, cstart, ccount, td->d_start[level],
                                   td->d_end[level], level, t->D, N,
                                   lap_plane(t, level, 0, c), lap_plane(t, level, 1, c),
                                   od->m2l_plan_dev[N], od->m2l_A1[N], od->m2l_A3[N],
                                   m2l_coax(od, N, t->D / (double)(1u << level)),
                                   od->m2l_CS[N], od->m2l_plan[N]));
    wbfmm_hip_free(cls);
  } else {
    unsigned c;
    ensure_lap_coax(od);
    for (c = 0; c < t->nq; c++)
      HIPOK(wbfmm_hip_lap_m2l(tl, nt, td->d_start[level], td->d_end[level], level,
                              t->D, t->order_s[level], t->order_r[level],
                              lap_plane(t, level, 0, c), lap_plane(t, level, 1, c),
                              od->d_rot, &od->lay, od->d_lap_coax));
  }
  if (level == t->depth) return 0;
  if (!t->mpr[level + 1]) WBFMM_FATAL("coefficients not initialised at level %u", level + 1);
  if ((int)t->order_r[level + 1] > od->lay.L) WBFMM_FATAL("operators built to order %d", od->lay.L);
  {
    unsigned c;
    for (c = 0; c < t->nq; c++)
      lap_l2l_plane(t, od, level, tl, nt, lap_plane(t, level, 1, c),
                    lap_plane(t, level + 1, 1, c));
  }
  return 0;
}

gint wbfmm_laplace_downward_pass_ref(wbfmm_tree_t *t,
                                     wbfmm_shift_operators_t *op,
                                     guint level, gdouble *work,
                                     gint nthreads)
{
  (void)work;
  (void)nthreads;
  return lap_downward(t, op, level);
}

gint wbfmm_laplace_downward_pass_avx(wbfmm_tree_t *t,
                                     wbfmm_shift_operators_t *op,
                                     guint level, gdouble *work,
                                     gint nthreads)
{
  (void)work;
  (void)nthreads;
  return lap_downward(t, op, level);
}

/* stage n points (host or device, byte stride) as 4-double device rows
   with identity ids; returns device buffers that the caller frees */
static void stage_targets(const void *xt, size_t tstr, unsigned n,
                          double **dx, uint32_t **did, uint64_t **dcodes,
                          const double *x0, double D, int sort_them)
{
  unsigned nbad = 0;
  double *raw;
  HIPOK(wbfmm_hip_malloc((void **)&raw, sizeof(double) * 4 * (n ? n : 1)));
  HIPOK(wbfmm_hip_malloc((void **)dx, sizeof(double) * 4 * (n ? n : 1)));
  HIPOK(wbfmm_hip_malloc((void **)did, sizeof(uint32_t) * (n ? n : 1)));
  HIPOK(wbfmm_hip_malloc((void **)dcodes, sizeof(uint64_t) * (n ? n : 1)));
  HIPOK(wbfmm_hip_morton(xt, tstr, n, x0, D, raw, *dcodes, *did, &nbad));
  if (nbad)
    WBFMM_FATAL("%u target points do not lie in the tree's domain", nbad);
  if (sort_them) HIPOK(wbfmm_hip_sort(*dcodes, *did, n));
  HIPOK(wbfmm_hip_gather4(raw, *did, n, *dx));
  HIPOK(wbfmm_hip_sync());
  wbfmm_hip_free(raw);
}

/* device view of a caller field array: f itself when it is device memory,
   else a staged copy that is written back by field_finish */
static double *field_begin(double *f, size_t nelem, int *staged)
{
  double *d;
  if (wbfmm_hip_is_device(f)) { *staged = 0; return f; }
  HIPOK(wbfmm_hip_malloc((void **)&d, sizeof(double) * (nelem ? nelem : 1)));
  HIPOK(wbfmm_hip_memcpy(d, f, sizeof(double) * nelem));
  *staged = 1;
  return d;
}

static void field_finish(double *f, double *d, size_t nelem, int staged)
{
  if (!staged) return;
  HIPOK(wbfmm_hip_memcpy(f, d, sizeof(double) * nelem));
  wbfmm_hip_free(d);
}

/* values per target of a Laplace field request with nq components:
   field bit 1 the potentials (nq), bit 2 the gradients (3 nq), or bit 4
   the curl (3) and, with bit 2, its gradient (9 more) */
static int lap_vec_width(guint field, int nq)
{
  if (field & WBFMM_FIELD_CURL) return (field & WBFMM_FIELD_GRADIENT) ? 12 : 3;
  return ((field & 1) ? nq : 0) + ((field & 2) ? 3 * nq : 0);
}

static void lap_field_run(tree_dev *td, const double *dxt,
                          const uint32_t *dtid, unsigned n, double *src,
                          gint sstr, double *d, gint dstr, int field,
                          double *f, gint fstr, size_t nf_rows)
{
  wbfmm_tree_t *t = &td->t;
  unsigned lv = t->depth;
  double *q, *dp, *df;
  int staged;
  size_t nelem;
  if (!t->mpr[lv]) WBFMM_FATAL("coefficients not initialised at level %u", lv);
  if (t->normals == NULL && d != NULL)
    WBFMM_FATAL("no normals in tree but dipole strengths specified (d != NULL)");
  if (t->nq > 1 || (field & 4)) {
    /* vector sources: nq strengths per point, the potentials and
       gradients of every component, or the curl (and its gradient) */
    const int width = lap_vec_width(field, (int)t->nq);
    if ((field & 4) && d) WBFMM_FATAL("the curl is defined for monopole sources only");
    q = gather_strengths_n(td, src, sstr, 0, (int)t->nq);
    dp = gather_strengths_n(td, d, dstr, 1, (int)t->nq);
    nelem = nf_rows ? (nf_rows - 1) * (size_t)fstr + width : 0;
    df = field_begin(f, nelem, &staged);
    HIPOK(wbfmm_hip_lap_field_vec(dxt, dtid, n, td->d_start[lv], td->d_end[lv], td->d_xyz, q,
                                  dp ? td->d_nrm : NULL, dp, (int)t->nq, lv, ORIGIN(t), t->D,
                                  t->order_r[lv], (const double *)t->mpr[lv],
                                  ((size_t)1 << (3 * lv)) * (t->order_r[lv] + 1) *
                                      (t->order_r[lv] + 1),
                                  field, df, fstr));
    field_finish(f, df, nelem, staged);
    return;
  }
  q = gather_strengths(td, src, sstr, 0);
  dp = gather_strengths(td, d, dstr, 1);
  nelem = nf_rows ? (nf_rows - 1) * (size_t)fstr + ((field & 1) ? 1 : 0) + ((field & 2) ? 3 : 0) : 0;
  df = field_begin(f, nelem, &staged);
  if (td->prec32)
    HIPOK(wbfmm_hip_lap_field_f32(dxt, dtid, n, td->d_start[lv], td->d_end[lv],
                                  td->d_xyz, q, dp ? td->d_nrm : NULL, dp, lv,
                                  ORIGIN(t), t->D, t->order_r[lv],
                                  (const float *)t->mpr[lv], field, df, fstr));
  else
    HIPOK(wbfmm_hip_lap_field(dxt, dtid, n, td->d_start[lv], td->d_end[lv],
                              td->d_xyz, q, dp ? td->d_nrm : NULL, dp, lv,
                              ORIGIN(t), t->D, t->order_r[lv],
                              (double *)t->mpr[lv], field, df, fstr));
  field_finish(f, df, nelem, staged);
}

gint wbfmm_laplace_field_targets(wbfmm_tree_t *t, gdouble *xt, gsize tstr,
                                 guint ntgt, gdouble *src, gint sstr,
                                 gdouble *d, gint dstr, guint field,
                                 gdouble *f, gint fstr)
{
  tree_dev *td = TD(t);
  double *dx;
  uint32_t *did;
  uint64_t *dc;
  int fl = field & 7;
  if (t->problem != WBFMM_PROBLEM_LAPLACE) return -2;
  if (fl == 0 || (field & ~7u)) return -2;
  if ((fl & 4) && t->nq < 3) return -2;
  if (fstr < lap_vec_width(fl, (int)t->nq)) return -2;
  if (ntgt == 0) return 0;
  stage_targets(xt, tstr, ntgt, &dx, &did, &dc, ORIGIN(t),
                t->D, 1);
  lap_field_run(td, dx, did, ntgt, src, sstr, d, dstr, fl, f, fstr, ntgt);
  HIPOK(wbfmm_hip_sync());
  wbfmm_hip_free(dx);
  wbfmm_hip_free(did);
  wbfmm_hip_free(dc);
  return 0;
}

gint wbfmm_laplace_box_field(wbfmm_tree_t *t, guint level, guint b,
                             gdouble *src, gint sstr, gdouble *d, gint dstr,
                             guint field, gboolean eval_neighbours,
                             gdouble *x, gdouble *f, gint fstr,
                             gdouble *work)
{
  tree_dev *td = TD(t);
  double *dx;
  uint32_t *did;
  uint64_t *dc;
  int fl;
  (void)work;
  (void)b;
  if (t->problem != WBFMM_PROBLEM_LAPLACE) WBFMM_FATAL("tree is not a Laplace tree");
  if (level != t->depth) WBFMM_FATAL("field evaluation is at the leaf level only");
  /* the field types of src/local-laplace.c:1293-1336 */
  switch (field) {
    case WBFMM_FIELD_POTENTIAL: fl = 1; break;
    case WBFMM_FIELD_GRADIENT:
    case WBFMM_FIELD_GRADIENT | WBFMM_FIELD_POTENTIAL: fl = 2; break;
    case WBFMM_FIELD_CURL:
    case WBFMM_FIELD_CURL | WBFMM_FIELD_POTENTIAL: fl = 4; break;
    case WBFMM_FIELD_CURL | WBFMM_FIELD_GRADIENT:
    case WBFMM_FIELD_CURL | WBFMM_FIELD_GRADIENT | WBFMM_FIELD_POTENTIAL: fl = 6; break;
    default: WBFMM_FATAL("unrecognized field type %u", field);
  }
  if (fl == 2 && fstr < 3 && t->nq != 1)
    WBFMM_FATAL("field data stride (%d) must be greater than two", fstr);
  if ((fl & 4) && t->nq < 3)
    WBFMM_FATAL("not enough source components (%u) for curl calculation", t->nq);
  stage_targets(x, 3 * sizeof(double), 1, &dx, &did, &dc,
                ORIGIN(t), t->D, 0);
  if (!eval_neighbours) { src = NULL; d = NULL; }
  {
    /* potential: component k at f[k fstr]; gradient: d_j of component k
       at f[k fstr + j]; curl: f[0..2]; curl and gradient: f[0..11] */
    double tmp[4 * LAP_MAX_NQ];
    const int width = lap_vec_width(fl, (int)t->nq);
    int k, j;
    memset(tmp, 0, sizeof(tmp));
    lap_field_run(td, dx, did, 1, src, sstr, d, dstr, fl, tmp, width, 1);
    if (fl == 1)
      for (k = 0; k < (int)t->nq; k++) f[k * fstr] += tmp[k];
    else if (fl == 2)
      for (k = 0; k < (int)t->nq; k++)
        for (j = 0; j < 3; j++) f[k * fstr + j] += tmp[3 * k + j];
    else
      for (k = 0; k < width; k++) f[k] += tmp[k];
  }
  wbfmm_hip_free(dx);
  wbfmm_hip_free(did);
  wbfmm_hip_free(dc);
  return 0;
}

/* the direct sums with vector sources (src/direct-laplace.c:87-198,
   :200-336): potential f[j fstr], gradient f[j fstr + c] of component j;
   curl f[0..2] and its gradient f[3..11] (d_c u_i at 3 + 3c + i), the
   curl terms skipping |x - y| <= 1e-6 as the reference's do */
static gint lap_direct_vec(gdouble *xs, gint xstride, gdouble *n, gint nstr, gint nsrc,
                           gdouble *src, gint sstr, gdouble *d, gint dstr, gint nq, int fl,
                           gdouble *xf, gdouble *f, gint fstr)
{
  double *dxs, *dq = NULL, *dd = NULL, *dn = NULL, *dxt, *df, tmp[4 * LAP_MAX_NQ];
  uint32_t *ids;
  const int width = lap_vec_width((guint)fl, nq);
  int j, c;
  HIPOK(wbfmm_hip_malloc((void **)&dxs, sizeof(double) * 4 * (nsrc ? nsrc : 1)));
  HIPOK(wbfmm_hip_stage4(xs, xstride * sizeof(double), nsrc, dxs));
  HIPOK(wbfmm_hip_malloc((void **)&ids, sizeof(uint32_t) * (nsrc ? nsrc : 1)));
  {
    uint32_t *h = malloc(sizeof(uint32_t) * (nsrc ? nsrc : 1));
    gint i2;
    if (!h) WBFMM_FATAL("out of host memory");
    for (i2 = 0; i2 < nsrc; i2++) h[i2] = i2;
    HIPOK(wbfmm_hip_memcpy(ids, h, sizeof(uint32_t) * nsrc));
    free(h);
  }
  if (src) {
    HIPOK(wbfmm_hip_malloc((void **)&dq, sizeof(double) * nq * (nsrc ? nsrc : 1)));
    HIPOK(wbfmm_hip_gather_strided(src, sstr, nq, ids, nsrc, dq));
  }
  if (d) {
    HIPOK(wbfmm_hip_malloc((void **)&dd, sizeof(double) * nq * (nsrc ? nsrc : 1)));
    HIPOK(wbfmm_hip_gather_strided(d, dstr, nq, ids, nsrc, dd));
    HIPOK(wbfmm_hip_malloc((void **)&dn, sizeof(double) * 4 * (nsrc ? nsrc : 1)));
    HIPOK(wbfmm_hip_stage4(n, nstr * sizeof(double), nsrc, dn));
  }
  HIPOK(wbfmm_hip_malloc((void **)&dxt, sizeof(double) * 4));
  {
    double h[4] = {xf[0], xf[1], xf[2], 0.0};
    HIPOK(wbfmm_hip_memcpy(dxt, h, sizeof(h)));
  }
  HIPOK(wbfmm_hip_malloc((void **)&df, sizeof(double) * width));
  HIPOK(wbfmm_hip_memset(df, 0, sizeof(double) * width));
  HIPOK(wbfmm_hip_lap_direct_vec(dxs, dq, dn, dd, nq, (unsigned)nsrc, dxt, 1, fl,
                                 (fl & 4) ? 1e-6 : 0.0, df, width));
  HIPOK(wbfmm_hip_memcpy(tmp, df, sizeof(double) * width));
  if (fl == 1)
    for (j = 0; j < nq; j++) f[j * fstr] += tmp[j];
  else if (fl == 2)
    for (j = 0; j < nq; j++)
      for (c = 0; c < 3; c++) f[j * fstr + c] += tmp[3 * j + c];
  else
    for (j = 0; j < width; j++) f[j] += tmp[j];
  wbfmm_hip_free(dxs); wbfmm_hip_free(ids);
  wbfmm_hip_free(dq); wbfmm_hip_free(dd); wbfmm_hip_free(dn);
  wbfmm_hip_free(dxt); wbfmm_hip_free(df);
  return 0;
}

gint wbfmm_laplace_field_direct(gdouble *xs, gint xstride, gdouble *n,
                                gint nstr, gint nsrc, gdouble *src,
                                gint sstr, gdouble *d, gint dstr, gint nq,
                                guint field, gdouble *xf, gdouble *f,
                                gint fstr)
{
  /* FP64 direct sum on the GPU for one field point, as the reference's
     wbfmm_laplace_field_direct (src/direct-laplace.c:274-336) */
  double *dxs, *dq = NULL, *dd = NULL, *dn = NULL, *dxt, *df;
  uint32_t *ids = NULL;
  int fl, i;
  double tmp[4] = {0, 0, 0, 0};
  if (src == NULL && d == NULL) return 0;
  if (n != NULL && d == NULL)
    WBFMM_FATAL("normals specified but no dipole strengths (d == NULL)");
  switch (field) {
    case WBFMM_FIELD_POTENTIAL: fl = 1; break;
    case WBFMM_FIELD_GRADIENT:
    case WBFMM_FIELD_GRADIENT | WBFMM_FIELD_POTENTIAL: fl = 2; break;
    case WBFMM_FIELD_CURL:
    case WBFMM_FIELD_CURL | WBFMM_FIELD_POTENTIAL: fl = 4; break;
    case WBFMM_FIELD_CURL | WBFMM_FIELD_GRADIENT:
    case WBFMM_FIELD_CURL | WBFMM_FIELD_GRADIENT | WBFMM_FIELD_POTENTIAL: fl = 6; break;
    default: WBFMM_FATAL("unrecognized field type %u", field);
  }
  if (nq < 1 || nq > LAP_MAX_NQ)
    WBFMM_FATAL("%d source components (1 to %d are supported)", nq, LAP_MAX_NQ);
  if ((fl & 4) && nq < 3)
    WBFMM_FATAL("not enough source components (%d) for curl calculation", nq);
  if (fl == 2 && fstr < 3 && nq > 1)
    WBFMM_FATAL("field data stride (%d) must be greater than two", fstr);
  if ((fl & 4) && d) WBFMM_FATAL("the curl is defined for monopole sources only");
  if (nq > 1 || (fl & 4))
    return lap_direct_vec(xs, xstride, n, nstr, nsrc, src, sstr, d, dstr, nq, fl, xf, f, fstr);
  HIPOK(wbfmm_hip_malloc((void **)&dxs, sizeof(double) * 4 * (nsrc ? nsrc : 1)));
  HIPOK(wbfmm_hip_stage4(xs, xstride * sizeof(double), nsrc, dxs));
  if (src) {
    HIPOK(wbfmm_hip_malloc((void **)&dq, sizeof(double) * (nsrc ? nsrc : 1)));
    HIPOK(wbfmm_hip_malloc((void **)&ids, sizeof(uint32_t) * (nsrc ? nsrc : 1)));
  }
  if (src || d) {
    uint32_t *h = malloc(sizeof(uint32_t) * (nsrc ? nsrc : 1));
    gint i2;
    for (i2 = 0; i2 < nsrc; i2++) h[i2] = i2;
    if (!ids) HIPOK(wbfmm_hip_malloc((void **)&ids, sizeof(uint32_t) * (nsrc ? nsrc : 1)));
    HIPOK(wbfmm_hip_memcpy(ids, h, sizeof(uint32_t) * nsrc));
    free(h);
  }
  if (src) HIPOK(wbfmm_hip_gather_strided(src, sstr, 1, ids, nsrc, dq));
  if (d) {
    HIPOK(wbfmm_hip_malloc((void **)&dd, sizeof(double) * (nsrc ? nsrc : 1)));
    HIPOK(wbfmm_hip_gather_strided(d, dstr, 1, ids, nsrc, dd));
    HIPOK(wbfmm_hip_malloc((void **)&dn, sizeof(double) * 4 * (nsrc ? nsrc : 1)));
    HIPOK(wbfmm_hip_stage4(n, nstr * sizeof(double), nsrc, dn));
  }
  HIPOK(wbfmm_hip_malloc((void **)&dxt, sizeof(double) * 4));
  {
    double h[4] = {xf[0], xf[1], xf[2], 0.0};
    HIPOK(wbfmm_hip_memcpy(dxt, h, sizeof(h)));
  }
  HIPOK(wbfmm_hip_malloc((void **)&df, sizeof(double) * 4));
  HIPOK(wbfmm_hip_memset(df, 0, sizeof(double) * 4));
  HIPOK(wbfmm_hip_lap_direct(dxs, dq, dn, dd, nsrc, dxt, 1, fl, 0.0, df, 4));
  HIPOK(wbfmm_hip_memcpy(tmp, df, sizeof(double) * 4));
  if (fl == 1) f[0] += tmp[0];
  else for (i = 0; i < 3; i++) f[i] += tmp[i];
  wbfmm_hip_free(dxs); wbfmm_hip_free(ids);
  wbfmm_hip_free(dq); wbfmm_hip_free(dd); wbfmm_hip_free(dn);
  wbfmm_hip_free(dxt); wbfmm_hip_free(df);
  return 0;
}

/* ------------------------------------------------------------------ */
/* target lists                                                       */
/* ------------------------------------------------------------------ */

wbfmm_target_list_t *wbfmm_target_list_new(wbfmm_tree_t *t, guint npts)
{
  tlist_dev *ld = calloc(1, sizeof(tlist_dev));
  wbfmm_target_list_t *l = &ld->l;
  l->t = t;
  l->maxpoints = npts;
  l->npoints = 0;
  l->size = sizeof(double);
  l->complex = (t->problem == WBFMM_PROBLEM_HELMHOLTZ);
  return l;
}

/* mark the boxes holding targets at every level and rebuild the active
   lists (boxes with sources or targets) the downward pass runs over */
static void register_targets(tree_dev *td, const uint64_t *dcodes, unsigned n)
{
  unsigned l;
  for (l = 0; l <= td->t.depth; l++) {
    size_t nb = (size_t)1 << (3 * l);
    if (!td->d_flag[l]) {
      HIPOK(wbfmm_hip_malloc((void **)&td->d_flag[l], sizeof(uint32_t) * nb));
      HIPOK(wbfmm_hip_malloc((void **)&td->d_act[l], sizeof(uint32_t) * nb));
    }
    HIPOK(wbfmm_hip_memset(td->d_flag[l], 0, sizeof(uint32_t) * nb));
    HIPOK(wbfmm_hip_mark_boxes(td->d_codes, td->t.npoints, l, td->d_flag[l]));
    HIPOK(wbfmm_hip_mark_boxes(dcodes, n, l, td->d_flag[l]));
    HIPOK(wbfmm_hip_compact_boxes(td->d_start[l], td->d_end[l], td->d_flag[l],
                                  l, td->d_act[l], &td->nact[l]));
  }
  td->have_targets = 1;
}

gint wbfmm_target_list_add_points(wbfmm_target_list_t *l, gpointer pts,
                                  gsize pstr, guint npts)
{
  tlist_dev *ld = (tlist_dev *)l;
  tree_dev *td = TD(l->t);
  if (npts > l->maxpoints)
    WBFMM_FATAL("too many points (%u) for target list (%u)", npts, l->maxpoints);
  if (ld->d_xt) {
    wbfmm_hip_free(ld->d_xt);
    wbfmm_hip_free(ld->d_tid);
    wbfmm_hip_free(ld->d_codes);
  }
  l->points = (char *)pts;
  l->pstr = pstr;
  l->npoints = npts;
  stage_targets(pts, pstr, npts, &ld->d_xt, &ld->d_tid, &ld->d_codes,
                ORIGIN(l->t), l->t->D, 1);
  register_targets(td, ld->d_codes, npts);
  return 0;
}

gint wbfmm_target_list_coefficients_init(wbfmm_target_list_t *l, guint field)
{
  l->field = field;
  return 0;
}

gint wbfmm_laplace_target_list_local_coefficients(wbfmm_target_list_t *l,
                                                  guint source, gdouble *work)
{
  (void)work;
  l->source = source;
  if (l->t->problem != WBFMM_PROBLEM_LAPLACE) WBFMM_FATAL("tree is not a Laplace tree");
  if (l->t->nq != 1)
    WBFMM_FATAL("target lists serve one source component (nq = %u)", l->t->nq);
  return 0;
}

/* src/targets.c:424-603 (Helmholtz precompute): the GPU evaluates the
   regular functions on the fly, so only the wavenumber is kept */
gint wbfmm_target_list_local_coefficients(wbfmm_target_list_t *l, gdouble k,
                                          gdouble *work)
{
  (void)work;
  if (l->t->problem != WBFMM_PROBLEM_HELMHOLTZ) WBFMM_FATAL("tree is not a Helmholtz tree");
  ((tlist_dev *)l)->hel_k = k;
  return 0;
}

static void hel_field_run(tree_dev *td, double k, const double *dxt,
                          const uint32_t *dtid, const uint64_t *dcodes, unsigned n,
                          double *src, gint sstr, double *d, gint dstr, int fl, double *f,
                          gint fstr, size_t nf_rows);

/* fused L2P + near field for every target of the list.  Laplace: f holds
   1 (potential) or 3 (gradient) values, 4 for both (potential first);
   Helmholtz: 2, 6 or 8 doubles (complex values).  The reference's target
   lists stop at gradients (src/targets.c:280, :562); here both kinds
   evaluate them. */
gint wbfmm_target_list_local_field(wbfmm_target_list_t *l, gdouble *src,
                                   gint sstr, gdouble *d, gint dstr,
                                   gdouble *f, gint fstr)
{
  tlist_dev *ld = (tlist_dev *)l;
  tree_dev *td = TD(l->t);
  int fl = (int)(l->field & 3);
  if (fl == 0) WBFMM_FATAL("unsupported field type %u", l->field);
  if (l->t->problem == WBFMM_PROBLEM_HELMHOLTZ) {
    if (fstr < ((fl & 1) ? 2 : 0) + ((fl & 2) ? 6 : 0))
      WBFMM_FATAL("field stride %d too small", fstr);
    hel_field_run(td, ld->hel_k, ld->d_xt, ld->d_tid, ld->d_codes, l->npoints, src, sstr, d,
                  dstr, fl, f, fstr, l->npoints);
    return 0;
  }
  if (fstr < ((fl & 1) ? 1 : 0) + ((fl & 2) ? 3 : 0))
    WBFMM_FATAL("field stride %d too small", fstr);
  lap_field_run(td, ld->d_xt, ld->d_tid, l->npoints, src, sstr, d, dstr, fl,
                f, fstr, l->npoints);
  return 0;
}

gint wbfmm_target_list_free(wbfmm_target_list_t *l)
{
  tlist_dev *ld = (tlist_dev *)l;
  if (!l) return 0;
  wbfmm_hip_free(ld->d_xt);
  wbfmm_hip_free(ld->d_tid);
  wbfmm_hip_free(ld->d_codes);
  free(ld);
  return 0;
}

gint wbfmm_tree_free(wbfmm_tree_t *t)
{
  tree_dev *td = TD(t);
  int l;
  if (!t) return 0;
  wbfmm_hip_sync();
  for (l = 0; l <= WBFMM_TREE_MAX_DEPTH; l++) {
    wbfmm_hip_free(td->d_start[l]);
    wbfmm_hip_free(td->d_end[l]);
    wbfmm_hip_free(td->d_occ[l]);
    wbfmm_hip_free(td->d_flag[l]);
    wbfmm_hip_free(td->d_act[l]);
    wbfmm_hip_free(t->mps[l]);
    wbfmm_hip_free(t->mpr[l]);
    free(t->boxes[l]);
    free(td->hsc[l]);
    wbfmm_hip_free(td->d_hsc[l]);
  }
  wbfmm_hip_free(td->d_xyz);
  wbfmm_hip_free(td->d_xyz_raw);
  wbfmm_hip_free(td->d_nrm);
  wbfmm_hip_free(td->d_codes);
  wbfmm_hip_free(td->d_idx);
  wbfmm_hip_free(td->d_q);
  wbfmm_hip_free(td->d_dip);
  free(t->ip);
  free(td);
  return 0;
}

gint wbfmm_hip_synchronize(void) { return wbfmm_hip_sync(); }

/* ------------------------------------------------------------------ */
/* Helmholtz path                                                     */
/* ------------------------------------------------------------------ */

/* src/boxes.c:536-583: complex coefficient arrays of level l, (n+1)^2
   complex per box (box-major on the device) */
gint wbfmm_tree_coefficient_init(wbfmm_tree_t *t, guint l, guint nr, guint ns)
{
  tree_dev *td = TD(t);
  size_t nb = (size_t)1 << (3 * l), es = td->prec32 ? sizeof(float) : sizeof(double);
  if ((int)t->nq < 1)
    WBFMM_FATAL("tree has invalid number of components in source terms (%d)", t->nq);
  if (t->nq != 1) WBFMM_FATAL("only nq == 1 is supported (nq = %u)", t->nq);
  if (t->problem != WBFMM_PROBLEM_HELMHOLTZ) WBFMM_FATAL("tree is not a Helmholtz tree");
  if (l > t->depth) WBFMM_FATAL("level %u below tree depth %u", l, t->depth);
  if (ns > WBFMM_HIP_MAX_ORDER || nr > WBFMM_HIP_MAX_ORDER)
    WBFMM_FATAL("order above %d", WBFMM_HIP_MAX_ORDER);
  if (td->prec32 && td->host_mirror)
    WBFMM_FATAL("FP32 coefficients cannot be mirrored to the host (wbfmm_tree_host_mirror)");
  t->order_s[l] = ns;
  t->order_r[l] = nr;
  /* mixed precision: FP32 storage until the wavenumber puts the level in
     FP64 (hel_prepare) */
  td->h64[l] = 0;
  td->hk_set = 0;
  if (ns) {
    alloc_level(td, l, 0, nb * 2 * (ns + 1) * (ns + 1) * es);
    if (td->host_mirror) set_box_pointers(td, l, 0, 2 * (ns + 1) * (ns + 1));
  }
  if (nr) {
    alloc_level(td, l, 1, nb * 2 * (nr + 1) * (nr + 1) * es);
    if (td->host_mirror) set_box_pointers(td, l, 1, 2 * (nr + 1) * (nr + 1));
  }
  return 0;
}

/* ---- mixed-precision Helmholtz ------------------------------------------
   FP32 levels hold M~_n = M_n / s_n and L~_n = L_n s_n with
     s_n = min(1, x^n / (2n+1)!!),  x = k a_l,  a_l = sqrt(3)/2 w_l,
   the small-argument size of j_n(k r) over the box (j_n(x) ~ x^n/(2n+1)!!
   for n > x): multipole coefficients of the level's boxes, and local
   coefficients times the regular functions at their targets, are then
   O(1) whatever the order, where unscaled FP32 would underflow (j_23(0.2)
   ~ 1e-45).  Every translation is the FP64 operator with its output degree
   n and input degree l scaled, rounded to FP32 once; rotations keep the
   degree and need no scaling.  Levels with k a_l above a threshold (the
   high-frequency levels near the root, default 2, WBFMM_HEL_FP64_KA) and
   levels whose order the FP32 matrix-core passes do not serve stay FP64;
   passes between an FP64 and an FP32 level run in FP64 on a converted
   copy.  Tables and recurrences stay FP64. */
static double hel_scale_arg(double k, double D, unsigned l)
{
  return k * 0.5 * sqrt(3.0) * D / (double)(1u << l);
}

static void hel_scales(double x, int nmax, double *sc)
{
  double lg = 0.0, lx = x > 1e-300 ? log(x) : log(1e-300);
  int n;
  for (n = 0; n <= nmax; n++) {
    lg += log(2.0 * n + 1.0);
    sc[n] = n * lx - lg < 0.0 ? exp(fmax(n * lx - lg, -640.0)) : 1.0;
  }
}

static double hel_fp64_ka(void)
{
  const char *e = getenv("WBFMM_HEL_FP64_KA");
  return e ? atof(e) : 2.0;
}

/* levels of a mixed-precision Helmholtz tree for wavenumber k: FP64 where
   k a_l exceeds the threshold or the orders are not served by the FP32
   passes (levels moved between precisions are reallocated, zeroed), and
   the scales of the FP32 levels */
static void hel_prepare(wbfmm_tree_t *t, double k)
{
  tree_dev *td = TD(t);
  unsigned l;
  if (!td->prec32 || (td->hk_set && td->hk == k)) return;
  for (l = 1; l <= t->depth; l++) {
    const unsigned ns = t->order_s[l], nr = t->order_r[l], N = ns > nr ? ns : nr;
    const size_t nb = (size_t)1 << (3 * l);
    const int want64 = hel_scale_arg(k, t->D, l) > hel_fp64_ka() || ns != nr ||
                       (int)N < M2L_MFMA_MIN || (int)N > HEL_MFMA_MAX ||
                       l > M2L_MFMA_MAX_LEVEL;
    if (want64 != td->h64[l]) {
      const size_t es = want64 ? sizeof(double) : sizeof(float);
      if (ns && t->mps[l]) alloc_level(td, l, 0, nb * 2 * (ns + 1) * (ns + 1) * es);
      if (nr && t->mpr[l]) alloc_level(td, l, 1, nb * 2 * (nr + 1) * (nr + 1) * es);
      td->h64[l] = want64;
    }
    free(td->hsc[l]);
    td->hsc[l] = malloc(sizeof(double) * (N + 2));
    if (!td->hsc[l]) WBFMM_FATAL("out of host memory");
    hel_scales(hel_scale_arg(k, t->D, l), (int)N + 1, td->hsc[l]);
    wbfmm_hip_free(td->d_hsc[l]);
    HIPOK(wbfmm_hip_malloc((void **)&td->d_hsc[l], sizeof(double) * (N + 2)));
    HIPOK(wbfmm_hip_memcpy(td->d_hsc[l], td->hsc[l], sizeof(double) * (N + 2)));
  }
  td->hk = k;
  td->hk_set = 1;
}

/* coefficients of `level' are FP32 */
static int level_f32(tree_dev *td, guint level)
{
  if (!td->prec32) return 0;
  if (td->t.problem == WBFMM_PROBLEM_HELMHOLTZ) return !td->h64[level];
  return 1;
}

/* src/boxes.c:341-490: P2M at the leaf level; src, dipoles: complex
   strengths (2 doubles) at element stride sstr, dstr */
gint wbfmm_tree_leaf_expansions(wbfmm_tree_t *t, gdouble k, gdouble *src,
                                gint sstr, gdouble *dipoles, gint dstr,
                                gboolean zero_expansions, gdouble *work)
{
  tree_dev *td = TD(t);
  unsigned d = t->depth;
  double *q, *dp;
  (void)work;
  if ((int)t->nq < 1)
    WBFMM_FATAL("tree has invalid number of components in source terms (%d)", t->nq);
  if (t->problem != WBFMM_PROBLEM_HELMHOLTZ) WBFMM_FATAL("tree is not a Helmholtz tree");
  if (!t->mps[d]) WBFMM_FATAL("coefficients not initialised at level %u", d);
  hel_prepare(t, k); /* mixed precision: the levels' precision follows k */
  if (zero_expansions) HIPOK(wbfmm_hip_memset(t->mps[d], 0, td->mps_bytes[d]));
  if (src == NULL && dipoles == NULL) return 0;
  if (t->normals == NULL && dipoles != NULL)
    WBFMM_FATAL("no normals in tree but dipole strengths specified (dipoles != NULL)");
  q = gather_strengths_n(td, src, sstr, 0, 2);
  dp = gather_strengths_n(td, dipoles, dstr, 1, 2);
  if (level_f32(td, d)) {
    HIPOK(wbfmm_hip_hel_p2m_f32(td->d_occ[d], td->nocc[d], td->d_start[d], td->d_end[d],
                                td->d_xyz, q, dp ? td->d_nrm : NULL, dp, d, ORIGIN(t), t->D,
                                t->order_s[d], k, td->d_hsc[d], (float *)t->mps[d]));
    return 0;
  }
  HIPOK(wbfmm_hip_hel_p2m(td->d_occ[d], td->nocc[d], td->d_start[d], td->d_end[d],
                          td->d_xyz, q, dp ? td->d_nrm : NULL, dp, d,
                          ORIGIN(t), t->D, t->order_s[d], k,
                          (double *)t->mps[d]));
  return 0;
}

static void hel_k_check(ops_dev *od, double k)
{
  if (od->hel_k_set && od->hel_k != k)
    WBFMM_FATAL("coaxial operators built for k = %g, not %g", od->hel_k, k);
  od->hel_k = k;
  od->hel_k_set = 1;
}

static size_t coax_entries(guint L) { return (size_t)(L + 1) * (L + 2) * (L + 3) / 6; }

/* src/shift.c:777-827: S|R tables of level `level' (box width D/2^level)
   at the 15 list-4 distances, order L, kept on the device */
gint wbfmm_shift_operators_coaxial_SR_init(wbfmm_shift_operators_t *w, gdouble D,
                                           guint level, guint L, gdouble k,
                                           gdouble *work)
{
  ops_dev *od = (ops_dev *)w;
  double thetas[64], chi[343], rr[16], wb = D / (double)(1u << level), *h;
  int th_idx[343], rix[343], nth, nr, i;
  size_t ne = coax_entries(L);
  (void)work;
  if (level > WBFMM_TREE_MAX_DEPTH) WBFMM_FATAL("level %u too deep", level);
  if (L > w->Lmax) WBFMM_FATAL("order %u above the operators' %u", L, w->Lmax);
  hel_k_check(od, k);
  wbfmm_hip_offset_geometry(thetas, &nth, th_idx, chi, rr, &nr, rix);
  h = malloc(sizeof(double) * 2 * ne * nr);
  if (!h) WBFMM_FATAL("out of host memory");
  {
    double kr[16];
    for (i = 0; i < nr; i++) kr[i] = k * wb * rr[i];
    HIPOK(wbfmm_hip_hel_coaxial_tables(1, (int)L, nr, kr, h));
  }
  if (w->SR[level]) wbfmm_hip_free(w->SR[level]);
  HIPOK(wbfmm_hip_malloc(&w->SR[level], sizeof(double) * 2 * ne * nr));
  HIPOK(wbfmm_hip_memcpy(w->SR[level], h, sizeof(double) * 2 * ne * nr));
  /* the host copy stays for the matrix-core tiles, made on first use */
  free(od->sr_host[level]);
  od->sr_host[level] = h;
  memcpy(od->sr_rr, rr, sizeof(double) * nr);
  od->sr_nr = nr;
  wbfmm_hip_free(od->hel_A2[level]);
  od->hel_A2[level] = NULL;
  w->L[level] = L;
  return 0;
}

/* src/shift.c:829-861: R|R (= S|S) table for shifts between level
   `level' and its parents, distance sqrt(3)/2 D/2^level, order L */
gint wbfmm_shift_operators_coaxial_SS_init(wbfmm_shift_operators_t *w, gdouble D,
                                           guint level, guint L, gdouble k,
                                           gdouble *work)
{
  ops_dev *od = (ops_dev *)w;
  double wb = D / (double)(1u << level), *h;
  size_t ne = coax_entries(L);
  (void)work;
  if (level > WBFMM_TREE_MAX_DEPTH) WBFMM_FATAL("level %u too deep", level);
  if (L > w->Lmax) WBFMM_FATAL("order %u above the operators' %u", L, w->Lmax);
  hel_k_check(od, k);
  h = malloc(sizeof(double) * 2 * ne);
  if (!h) WBFMM_FATAL("out of host memory");
  HIPOK(wbfmm_hip_hel_coaxial_table(0, (int)L, k * wb * 0.5 * sqrt(3.0), h));
  if (w->SS[level]) wbfmm_hip_free(w->SS[level]);
  HIPOK(wbfmm_hip_malloc(&w->SS[level], sizeof(double) * 2 * ne));
  HIPOK(wbfmm_hip_memcpy(w->SS[level], h, sizeof(double) * 2 * ne));
  /* the host copy stays for the matrix-core tiles, made on first use */
  free(od->ss_host[level]);
  od->ss_host[level] = h;
  wbfmm_hip_free(od->hel_SS2[level]);
  od->hel_SS2[level] = NULL;
  od->ss_L[level] = L;
  return 0;
}

gint wbfmm_shift_operators_coaxial_reset(wbfmm_shift_operators_t *w)
{
  ops_dev *od = (ops_dev *)w;
  int n;
  for (n = 0; n <= WBFMM_TREE_MAX_DEPTH; n++) {
    wbfmm_hip_free(w->SR[n]);
    wbfmm_hip_free(w->SS[n]);
    w->SR[n] = w->SS[n] = NULL;
    w->L[n] = 0;
    od->ss_L[n] = 0;
    wbfmm_hip_free(od->hel_A2[n]);
    od->hel_A2[n] = NULL;
    free(od->sr_host[n]);
    od->sr_host[n] = NULL;
    wbfmm_hip_free(od->hel_SS2[n]);
    od->hel_SS2[n] = NULL;
    free(od->ss_host[n]);
    od->ss_host[n] = NULL;
    wbfmm_hip_free(od->f_hA2[n]);
    wbfmm_hip_free(od->f_hUp[n]);
    wbfmm_hip_free(od->f_hDn[n]);
    od->f_hA2[n] = od->f_hUp[n] = od->f_hDn[n] = NULL;
  }
  od->hel_k_set = 0;
  return 0;
}

/* H tables and offset geometry of the Helmholtz passes.  They depend on
   the order only, so one set per order is built for the whole process (by
   the first operator set that needs it, on host threads) and shared by
   every operator set of that order until the process ends. */
typedef struct {
  int state; /* 0 none, 1 being built, 2 ready */
  double *Hth, *Hud, *chi;
  int *th_idx, *rix;
  size_t nrot;
} hel_rot_set;
static hel_rot_set g_hel_rot[WBFMM_HIP_MAX_ORDER + 1];
static pthread_mutex_t g_hel_rot_mu = PTHREAD_MUTEX_INITIALIZER;
static pthread_cond_t g_hel_rot_cv = PTHREAD_COND_INITIALIZER;

static void ensure_hel_rotations(ops_dev *od)
{
  int L = (int)od->op.Lmax;
  hel_rot_set *e = &g_hel_rot[L];
  if (od->d_hel_Hth) return;
  mu_lock(&g_hel_rot_mu);
  while (e->state == 1) pthread_cond_wait(&g_hel_rot_cv, &g_hel_rot_mu);
  if (e->state == 0) {
    double thetas[64], chi[343], rr[16], tud[2];
    int th_idx[343], rix[343], nth, nr;
    hel_rot_set b;
    e->state = 1;
    mu_unlock(&g_hel_rot_mu);
    wbfmm_hip_offset_geometry(thetas, &nth, th_idx, chi, rr, &nr, rix);
    HIPOK(wbfmm_hip_H_tables(L, nth, thetas, &b.Hth, &b.nrot));
    tud[0] = acos(1.0 / sqrt(3.0));
    tud[1] = acos(-1.0 / sqrt(3.0));
    HIPOK(wbfmm_hip_H_tables(L, 2, tud, &b.Hud, &b.nrot));
    /* the copies from host memory return when done: the tables are ready
       for every stream */
    HIPOK(wbfmm_hip_malloc((void **)&b.th_idx, sizeof(th_idx)));
    HIPOK(wbfmm_hip_memcpy(b.th_idx, th_idx, sizeof(th_idx)));
    HIPOK(wbfmm_hip_malloc((void **)&b.rix, sizeof(rix)));
    HIPOK(wbfmm_hip_memcpy(b.rix, rix, sizeof(rix)));
    HIPOK(wbfmm_hip_malloc((void **)&b.chi, sizeof(chi)));
    HIPOK(wbfmm_hip_memcpy(b.chi, chi, sizeof(chi)));
    mu_lock(&g_hel_rot_mu);
    b.state = 2;
    *e = b;
    pthread_cond_broadcast(&g_hel_rot_cv);
  }
  od->d_hel_Hth = e->Hth;
  od->d_hel_Hud = e->Hud;
  od->hel_nrot = e->nrot;
  od->d_th_idx = e->th_idx;
  od->d_rix = e->rix;
  od->d_chi = e->chi;
  mu_unlock(&g_hel_rot_mu);
}

static int hel_shift_mfma(ops_dev *od, guint Nc, guint Np);
static void ensure_hel_shift_rot_tiles(ops_dev *od, int N);
static void ensure_hel_shift_coax_tiles(ops_dev *od, guint level, int N);

/* src/pass.c:581-632: M2M from level into level-1 */
/* FP64 M2M from child expansions src (level `level') into the parents'
   dst (level-1, overwritten) */
static void hel_m2m64(wbfmm_tree_t *t, ops_dev *od, guint level, const double *src, double *dst)
{
  tree_dev *td = TD(t);
  guint Nc = t->order_s[level], Np = t->order_s[level - 1];
  if (hel_shift_mfma(od, Nc, Np)) {
    int N = (int)Nc;
    ensure_plan(od, N);
    ensure_hel_shift_rot_tiles(od, N);
    ensure_hel_shift_coax_tiles(od, level, N);
    HIPOK(wbfmm_hip_hel_shift_mfma(0, td->d_occ[level - 1], td->nocc[level - 1],
                                   td->d_start[level], td->d_end[level], NULL, N, src, dst,
                                   od->m2l_plan_dev[N], od->hel_shA[N], od->hel_SS2[level],
                                   od->sh_CSup[N], od->m2l_plan[N]));
    return;
  }
  ensure_hel_rotations(od);
  HIPOK(wbfmm_hip_hel_m2m(td->d_occ[level - 1], td->nocc[level - 1], td->d_start[level],
                          td->d_end[level], (int)Nc, (int)Np, src, dst, od->d_hel_Hud,
                          od->hel_nrot, (double *)od->op.SS[level]));
}

static void hel_upward_mixed(wbfmm_tree_t *t, ops_dev *od, guint level);

gint wbfmm_upward_pass(wbfmm_tree_t *t, wbfmm_shift_operators_t *op, guint level,
                       gdouble *work)
{
  tree_dev *td = TD(t);
  ops_dev *od = (ops_dev *)op;
  guint Nc, Np;
  (void)work;
  if (level <= 1) WBFMM_FATAL("level %u <= 1", level);
  if (t->problem != WBFMM_PROBLEM_HELMHOLTZ) WBFMM_FATAL("tree is not a Helmholtz tree");
  if (!t->mps[level] || !t->mps[level - 1])
    WBFMM_FATAL("coefficients not initialised at levels %u, %u", level, level - 1);
  Nc = t->order_s[level];
  Np = t->order_s[level - 1];
  if (!op->SS[level]) WBFMM_FATAL("S|S operators not initialised at level %u", level);
  if ((Nc > Np ? Nc : Np) > od->ss_L[level] || (Nc > Np ? Nc : Np) > op->Lmax)
    WBFMM_FATAL("operators at level %u built to order %u", level, od->ss_L[level]);
  if (td->prec32) {
    hel_upward_mixed(t, od, level);
    return 0;
  }
  hel_m2m64(t, od, level, (const double *)t->mps[level], (double *)t->mps[level - 1]);
  return 0;
}

/* the matrix-core Helmholtz M2L is used for equal orders in
   M2L_MFMA_MIN..HEL_MFMA_MAX unless WBFMM_HEL_M2L=vector asks for the
   vector kernel (kept for comparison; the results agree to rounding) */
static int hel_mfma_enabled(void)
{
  static int on = -1;
  if (on < 0) {
    const char *e = getenv("WBFMM_HEL_M2L");
    on = !(e && strcmp(e, "vector") == 0);
  }
  return on;
}

/* rotation tiles of the Helmholtz matrix-core M2L of order N (both
   rotations), built once per process */
static void ensure_hel_rot_tiles(ops_dev *od, int N)
{
  if (od->hel_A[N]) return;
  mu_lock(&g_ord_mu[N]);
  if (!g_tab.hel_A[N]) {
    const wbfmm_m2l_plan_t *P = od->m2l_plan[N];
    size_t n = (size_t)P->NTH * P->KT * 64;
    double *h = malloc(sizeof(double) * n), *d;
    if (!h) WBFMM_FATAL("out of host memory");
    HIPOK(wbfmm_hip_hel_m2l_rot_tiles(P, h));
    HIPOK(wbfmm_hip_malloc((void **)&d, sizeof(double) * n));
    HIPOK(wbfmm_hip_memcpy(d, h, sizeof(double) * n));
    free(h);
    g_tab.hel_A[N] = d;
  }
  od->hel_A[N] = g_tab.hel_A[N];
  mu_unlock(&g_ord_mu[N]);
}

/* coaxial tiles of the Helmholtz matrix-core M2L at `level' for order N,
   from the level's host S|R tables (ordered as the plan's offset lengths) */
static void ensure_hel_coax_tiles(ops_dev *od, guint level, int N)
{
  const wbfmm_m2l_plan_t *P = od->m2l_plan[N];
  size_t ne = coax_entries(od->op.L[level]), n = (size_t)P->NR * 2 * P->KT * 64;
  double *E, *h;
  int i, j;
  if (od->hel_A2[level] && od->hel_A2_N[level] == N) return;
  if (!od->sr_host[level]) WBFMM_FATAL("S|R operators not initialised at level %u", level);
  if (od->sr_nr != P->NR) WBFMM_FATAL("%d offset lengths, the plan has %d", od->sr_nr, P->NR);
  E = malloc(sizeof(double) * 2 * ne * P->NR);
  h = malloc(sizeof(double) * n);
  if (!E || !h) WBFMM_FATAL("out of host memory");
  for (i = 0; i < P->NR; i++) {
    for (j = 0; j < od->sr_nr; j++)
      if (od->sr_rr[j] == P->rr[i]) break;
    if (j == od->sr_nr) WBFMM_FATAL("offset length %g not in the S|R tables", P->rr[i]);
    memcpy(E + (size_t)i * 2 * ne, od->sr_host[level] + (size_t)j * 2 * ne,
           sizeof(double) * 2 * ne);
  }
  HIPOK(wbfmm_hip_hel_m2l_coax_tiles(P, (int)od->op.L[level], P->NR, E, h));
  wbfmm_hip_free(od->hel_A2[level]);
  HIPOK(wbfmm_hip_malloc((void **)&od->hel_A2[level], sizeof(double) * n));
  HIPOK(wbfmm_hip_memcpy(od->hel_A2[level], h, sizeof(double) * n));
  od->hel_A2_N[level] = N;
  free(E);
  free(h);
}

/* rotation tiles of the Helmholtz matrix-core M2M / L2L of order N at the
   child polar angles acos(1/sqrt 3), acos(-1/sqrt 3), built once per
   process */
static void ensure_hel_shift_rot_tiles(ops_dev *od, int N)
{
  if (od->hel_shA[N]) return;
  mu_lock(&g_ord_mu[N]);
  if (!g_tab.hel_shA[N]) {
    const wbfmm_m2l_plan_t *P = od->m2l_plan[N];
    size_t n = (size_t)2 * P->KT * 64;
    double ths[2] = {acos(1.0 / sqrt(3.0)), acos(-1.0 / sqrt(3.0))}, *d;
    double *h = malloc(sizeof(double) * n);
    if (!h) WBFMM_FATAL("out of host memory");
    HIPOK(wbfmm_hip_hel_rot_tiles(P, 2, ths, h));
    HIPOK(wbfmm_hip_malloc((void **)&d, sizeof(double) * n));
    HIPOK(wbfmm_hip_memcpy(d, h, sizeof(double) * n));
    free(h);
    g_tab.hel_shA[N] = d;
  }
  od->hel_shA[N] = g_tab.hel_shA[N];
  mu_unlock(&g_ord_mu[N]);
}

/* real coaxial tiles of the Helmholtz matrix-core M2M / L2L between
   `level' and its parents, order N, from the level's host R|R table */
static void ensure_hel_shift_coax_tiles(ops_dev *od, guint level, int N)
{
  const wbfmm_m2l_plan_t *P = od->m2l_plan[N];
  size_t n = (size_t)P->KT * 64;
  double *h;
  if (od->hel_SS2[level] && od->hel_SS2_N[level] == N) return;
  if (!od->ss_host[level]) WBFMM_FATAL("S|S operators not initialised at level %u", level);
  h = malloc(sizeof(double) * 2 * n);
  if (!h) WBFMM_FATAL("out of host memory");
  HIPOK(wbfmm_hip_hel_m2l_coax_tiles(P, (int)od->ss_L[level], 1, od->ss_host[level], h));
  wbfmm_hip_free(od->hel_SS2[level]);
  HIPOK(wbfmm_hip_malloc((void **)&od->hel_SS2[level], sizeof(double) * n));
  HIPOK(wbfmm_hip_memcpy(od->hel_SS2[level], h, sizeof(double) * n));  /* real part */
  od->hel_SS2_N[level] = N;
  free(h);
}

/* the matrix-core M2M / L2L serves equal orders in M2L_MFMA_MIN..HEL_MFMA_MAX */
static int hel_shift_mfma(ops_dev *od, guint Nc, guint Np)
{
  (void)od;
  return Nc == Np && (int)Nc >= M2L_MFMA_MIN && (int)Nc <= HEL_MFMA_MAX && hel_mfma_enabled();
}

/* FP64 M2L at `level' for the nt target boxes tl */
static void hel_m2l64(wbfmm_tree_t *t, ops_dev *od, guint level, const uint32_t *tl, unsigned nt)
{
  tree_dev *td = TD(t);
  wbfmm_shift_operators_t *op = &od->op;
  guint Ns = t->order_s[level], Nr = t->order_r[level];
  if (Ns == Nr && (int)Ns >= M2L_MFMA_MIN && (int)Ns <= HEL_MFMA_MAX && hel_mfma_enabled() &&
      level <= M2L_MFMA_MAX_LEVEL) {
    /* matrix-core form: targets grouped by Morton child digit */
    int N = (int)Ns;
    unsigned cstart[8], ccount[8];
    uint32_t *cls;
    ensure_plan(od, N);
    ensure_hel_rot_tiles(od, N);
    ensure_hel_coax_tiles(od, level, N);
    HIPOK(wbfmm_hip_malloc((void **)&cls, sizeof(uint32_t) * (nt ? nt : 1)));
    HIPOK(wbfmm_hip_split_classes(tl, nt, cls, cstart, ccount));
    HIPOK(wbfmm_hip_hel_m2l_mfma(cls, cstart, ccount, td->d_start[level], td->d_end[level],
                                 level, N, (double *)t->mps[level], (double *)t->mpr[level],
                                 od->m2l_plan_dev[N], od->hel_A[N], od->hel_A2[level],
                                 od->m2l_CS[N], od->m2l_plan[N]));
    wbfmm_hip_free(cls);
  } else {
    ensure_hel_rotations(od); /* the vector kernels' H tables, on first use */
    HIPOK(wbfmm_hip_hel_m2l(tl, nt, td->d_start[level], td->d_end[level], level, (int)Ns,
                            (int)Nr, (double *)t->mps[level], (double *)t->mpr[level],
                            od->d_hel_Hth, od->hel_nrot, od->d_th_idx, od->d_chi, od->d_rix,
                            (double *)op->SR[level], coax_entries(op->L[level])));
  }
}

/* FP64 L2L from the parents tl (level `level', local expansions src) into
   their children's dst (level+1, added) */
static void hel_l2l64(wbfmm_tree_t *t, ops_dev *od, guint level, const uint32_t *tl,
                      unsigned nt, const double *src, double *dst)
{
  tree_dev *td = TD(t);
  guint Nr = t->order_r[level], Nc = t->order_r[level + 1];
  if (hel_shift_mfma(od, Nc, Nr)) {
    int N = (int)Nc;
    ensure_plan(od, N);
    ensure_hel_shift_rot_tiles(od, N);
    ensure_hel_shift_coax_tiles(od, level + 1, N);
    HIPOK(wbfmm_hip_hel_shift_mfma(1, tl, nt, NULL, NULL,
                                   td->have_targets ? td->d_flag[level + 1] : NULL, N, src, dst,
                                   od->m2l_plan_dev[N], od->hel_shA[N], od->hel_SS2[level + 1],
                                   od->sh_CSdn[N], od->m2l_plan[N]));
  } else {
    ensure_hel_rotations(od);
    HIPOK(wbfmm_hip_hel_l2l(tl, nt, td->have_targets ? td->d_flag[level + 1] : NULL, (int)Nr,
                            (int)Nc, src, dst, od->d_hel_Hud, od->hel_nrot,
                            (double *)od->op.SS[level + 1]));
  }
}

static void hel_downward_mixed(wbfmm_tree_t *t, ops_dev *od, guint level, const uint32_t *tl,
                               unsigned nt);

static gint hel_downward(wbfmm_tree_t *t, wbfmm_shift_operators_t *op, guint level)
{
  tree_dev *td = TD(t);
  ops_dev *od = (ops_dev *)op;
  uint32_t *tl;
  unsigned nt;
  guint Ns, Nr;
  if (level <= 1) WBFMM_FATAL("level %u <= 1", level);
  if (t->problem != WBFMM_PROBLEM_HELMHOLTZ) WBFMM_FATAL("tree is not a Helmholtz tree");
  if (!t->mpr[level] || !t->mps[level])
    WBFMM_FATAL("coefficients not initialised at level %u", level);
  Ns = t->order_s[level];
  Nr = t->order_r[level];
  if (!op->SR[level]) WBFMM_FATAL("S|R operators not initialised at level %u", level);
  if ((Ns > Nr ? Ns : Nr) > op->L[level])
    WBFMM_FATAL("S|R operators at level %u built to order %u", level, op->L[level]);
  if ((Ns > Nr ? Ns : Nr) > op->Lmax) WBFMM_FATAL("rotations built to order %u", op->Lmax);
  if (td->have_targets) {
    tl = td->d_act[level];
    nt = td->nact[level];
  } else {
    tl = all_boxes(level, &nt);
  }
  if (level < t->depth) {
    guint Nc = t->order_r[level + 1];
    if (!t->mpr[level + 1]) WBFMM_FATAL("coefficients not initialised at level %u", level + 1);
    if (!op->SS[level + 1]) WBFMM_FATAL("S|S operators not initialised at level %u", level + 1);
    if ((Nr > Nc ? Nr : Nc) > od->ss_L[level + 1] || (Nr > Nc ? Nr : Nc) > op->Lmax)
      WBFMM_FATAL("S|S operators at level %u built to order %u", level + 1, od->ss_L[level + 1]);
  }
  if (td->prec32) {
    hel_downward_mixed(t, od, level, tl, nt);
    return 0;
  }
  hel_m2l64(t, od, level, tl, nt);
  if (level == t->depth) return 0;
  hel_l2l64(t, od, level, tl, nt, (const double *)t->mpr[level], (double *)t->mpr[level + 1]);
  return 0;
}

/* ---- mixed-precision passes -------------------------------------------- */

/* the FP32 rotation tiles and phases of order N */
static void ensure_hel_f32_order(ops_dev *od, int N)
{
  const wbfmm_m2l_plan_t *P;
  ensure_plan(od, N);
  ensure_hel_rot_tiles(od, N);
  ensure_hel_shift_rot_tiles(od, N);
  P = od->m2l_plan[N];
  if (!od->f_hA[N]) od->f_hA[N] = dev_f32(od->hel_A[N], (size_t)P->NTH * P->KT * 64);
  if (!od->f_hshA[N]) od->f_hshA[N] = dev_f32(od->hel_shA[N], (size_t)2 * P->KT * 64);
  if (!od->f_hCS[N]) od->f_hCS[N] = dev_f32(od->m2l_CS[N], (size_t)343 * (N + 1) * 2);
  if (!od->f_hCSup[N]) od->f_hCSup[N] = dev_f32(od->sh_CSup[N], (size_t)8 * (N + 1) * 2);
  if (!od->f_hCSdn[N]) od->f_hCSdn[N] = dev_f32(od->sh_CSdn[N], (size_t)8 * (N + 1) * 2);
}

/* scaled FP32 coaxial tiles of the M2L at `level': L~ = s L, M = s M~ */
static void ensure_hel_f32_m2l_tiles(ops_dev *od, double D, guint level, int N, double k)
{
  const wbfmm_m2l_plan_t *P = od->m2l_plan[N];
  size_t ne = coax_entries(od->op.L[level]), n = (size_t)P->NR * 2 * P->KT * 64;
  double *E, *h, sc[WBFMM_HIP_MAX_ORDER + 2];
  int i, j;
  if (od->f_hA2[level] && od->f_hA2_N[level] == N && od->f_hA2_k[level] == k && od->f_hD == D)
    return;
  if (!od->sr_host[level]) WBFMM_FATAL("S|R operators not initialised at level %u", level);
  if (od->sr_nr != P->NR) WBFMM_FATAL("%d offset lengths, the plan has %d", od->sr_nr, P->NR);
  E = malloc(sizeof(double) * 2 * ne * P->NR);
  h = malloc(sizeof(double) * n);
  if (!E || !h) WBFMM_FATAL("out of host memory");
  for (i = 0; i < P->NR; i++) {
    for (j = 0; j < od->sr_nr; j++)
      if (od->sr_rr[j] == P->rr[i]) break;
    if (j == od->sr_nr) WBFMM_FATAL("offset length %g not in the S|R tables", P->rr[i]);
    memcpy(E + (size_t)i * 2 * ne, od->sr_host[level] + (size_t)j * 2 * ne,
           sizeof(double) * 2 * ne);
  }
  hel_scales(hel_scale_arg(k, D, level), N + 1, sc);
  HIPOK(wbfmm_hip_hel_coax_tiles_scaled(P, (int)od->op.L[level], P->NR, E, sc, sc, h));
  wbfmm_hip_free(od->f_hA2[level]);
  od->f_hA2[level] = host_f32(h, n);
  od->f_hA2_N[level] = N;
  od->f_hA2_k[level] = k;
  od->f_hD = D;
  free(E);
  free(h);
}

/* scaled FP32 coaxial tiles of the M2M from `level' into level-1 (up) and
   of the L2L from level-1 into `level' (down), from the level's R|R table */
static void ensure_hel_f32_shift_tiles(ops_dev *od, double D, guint level, int N, double k)
{
  const wbfmm_m2l_plan_t *P = od->m2l_plan[N];
  size_t n = (size_t)P->KT * 64;
  double *h, sc[WBFMM_HIP_MAX_ORDER + 2], sp[WBFMM_HIP_MAX_ORDER + 2];
  double ip[WBFMM_HIP_MAX_ORDER + 2];
  int i;
  if (od->f_hUp[level] && od->f_hUp_N[level] == N && od->f_hUp_k[level] == k &&
      od->f_hDn[level] && od->f_hDn_N[level] == N && od->f_hDn_k[level] == k && od->f_hD == D)
    return;
  if (!od->ss_host[level]) WBFMM_FATAL("S|S operators not initialised at level %u", level);
  h = malloc(sizeof(double) * 2 * n);
  if (!h) WBFMM_FATAL("out of host memory");
  hel_scales(hel_scale_arg(k, D, level), N + 1, sc);
  hel_scales(hel_scale_arg(k, D, level - 1), N + 1, sp);
  for (i = 0; i <= N + 1; i++) ip[i] = 1.0 / sp[i];
  /* M2M: M~_parent = M_parent / s_parent, M_child = s_child M~_child */
  HIPOK(wbfmm_hip_hel_coax_tiles_scaled(P, (int)od->ss_L[level], 1, od->ss_host[level], ip, sc, h));
  wbfmm_hip_free(od->f_hUp[level]);
  od->f_hUp[level] = host_f32(h, n); /* real part */
  /* L2L: L~_child = s_child L_child, L_parent = L~_parent / s_parent */
  HIPOK(wbfmm_hip_hel_coax_tiles_scaled(P, (int)od->ss_L[level], 1, od->ss_host[level], sc, ip, h));
  wbfmm_hip_free(od->f_hDn[level]);
  od->f_hDn[level] = host_f32(h, n);
  od->f_hUp_N[level] = od->f_hDn_N[level] = N;
  od->f_hUp_k[level] = od->f_hDn_k[level] = k;
  od->f_hD = D;
  free(h);
}

/* FP64 copy of an FP32 level (which: 0 mps, 1 mpr) for the listed boxes,
   unscaled; the rest of the copy is zero */
static double *hel_level_to64(wbfmm_tree_t *t, guint level, int which, const uint32_t *boxes,
                              unsigned nbox)
{
  tree_dev *td = TD(t);
  int N = (int)(which ? t->order_r[level] : t->order_s[level]), n;
  size_t bytes = ((size_t)1 << (3 * level)) * 2 * (N + 1) * (N + 1) * sizeof(double);
  double *d = NULL, us[WBFMM_HIP_MAX_ORDER + 2], *dus = NULL;
  HIPOK(wbfmm_hip_malloc((void **)&d, bytes));
  HIPOK(wbfmm_hip_memset(d, 0, bytes));
  for (n = 0; n <= N; n++) us[n] = which ? 1.0 / td->hsc[level][n] : td->hsc[level][n];
  HIPOK(wbfmm_hip_malloc((void **)&dus, sizeof(double) * (N + 1)));
  HIPOK(wbfmm_hip_memcpy(dus, us, sizeof(double) * (N + 1)));
  HIPOK(wbfmm_hip_hel_convert(0, 0, boxes, nbox, N, which ? t->mpr[level] : t->mps[level], dus, d));
  HIPOK(wbfmm_hip_free(dus));
  return d;
}

/* an FP64 level array d (unscaled) into an FP32 level, for the listed
   boxes: overwrite (add 0) or add (add 1) */
static void hel_level_to32(wbfmm_tree_t *t, guint level, int which, const uint32_t *boxes,
                           unsigned nbox, const double *d, int add)
{
  tree_dev *td = TD(t);
  int N = (int)(which ? t->order_r[level] : t->order_s[level]), n;
  double sc[WBFMM_HIP_MAX_ORDER + 2], *dsc = NULL;
  for (n = 0; n <= N; n++) sc[n] = which ? td->hsc[level][n] : 1.0 / td->hsc[level][n];
  HIPOK(wbfmm_hip_malloc((void **)&dsc, sizeof(double) * (N + 1)));
  HIPOK(wbfmm_hip_memcpy(dsc, sc, sizeof(double) * (N + 1)));
  HIPOK(wbfmm_hip_hel_convert(1, add, boxes, nbox, N, d, dsc, which ? t->mpr[level] : t->mps[level]));
  HIPOK(wbfmm_hip_free(dsc));
}

static void hel_require_k(wbfmm_tree_t *t, ops_dev *od)
{
  tree_dev *td = TD(t);
  if (!td->hk_set) WBFMM_FATAL("leaf expansions not formed (the wavenumber sets the precision)");
  if (od->hel_k_set && od->hel_k != td->hk)
    WBFMM_FATAL("operators for k = %g, tree formed for k = %g", od->hel_k, td->hk);
}

static void hel_upward_mixed(wbfmm_tree_t *t, ops_dev *od, guint level)
{
  tree_dev *td = TD(t);
  const int c32 = !td->h64[level], p32 = !td->h64[level - 1];
  const guint Nc = t->order_s[level], Np = t->order_s[level - 1];
  hel_require_k(t, od);
  if (c32 && p32 && Nc == Np) {
    int N = (int)Nc;
    ensure_hel_f32_order(od, N);
    ensure_hel_f32_shift_tiles(od, t->D, level, N, td->hk);
    HIPOK(wbfmm_hip_hel_shift_mfma_f32(0, td->d_occ[level - 1], td->nocc[level - 1],
                                       td->d_start[level], td->d_end[level], NULL, N,
                                       (const float *)t->mps[level], (float *)t->mps[level - 1],
                                       od->m2l_plan_dev[N], od->f_hshA[N], od->f_hUp[level],
                                       od->f_hCSup[N], od->m2l_plan[N]));
    return;
  }
  {
    /* a pass across precisions runs in FP64 on converted copies */
    double *src = c32 ? hel_level_to64(t, level, 0, td->d_occ[level], td->nocc[level])
                      : (double *)t->mps[level];
    double *dst = (double *)t->mps[level - 1];
    if (p32) {
      size_t bytes = ((size_t)1 << (3 * (level - 1))) * 2 * (Np + 1) * (Np + 1) * sizeof(double);
      HIPOK(wbfmm_hip_malloc((void **)&dst, bytes));
      HIPOK(wbfmm_hip_memset(dst, 0, bytes));
    }
    hel_m2m64(t, od, level, src, dst);
    if (p32) {
      hel_level_to32(t, level - 1, 0, td->d_occ[level - 1], td->nocc[level - 1], dst, 0);
      HIPOK(wbfmm_hip_free(dst));
    }
    if (c32) HIPOK(wbfmm_hip_free(src));
  }
}

static void hel_downward_mixed(wbfmm_tree_t *t, ops_dev *od, guint level, const uint32_t *tl,
                               unsigned nt)
{
  tree_dev *td = TD(t);
  const guint Ns = t->order_s[level], Nr = t->order_r[level];
  hel_require_k(t, od);
  if (!td->h64[level]) {
    int N = (int)Ns;
    unsigned cstart[8], ccount[8];
    uint32_t *cls;
    if (Ns != Nr) WBFMM_FATAL("FP32 level %u with unequal orders", level);
    ensure_hel_f32_order(od, N);
    ensure_hel_f32_m2l_tiles(od, t->D, level, N, td->hk);
    HIPOK(wbfmm_hip_malloc((void **)&cls, sizeof(uint32_t) * (nt ? nt : 1)));
    HIPOK(wbfmm_hip_split_classes(tl, nt, cls, cstart, ccount));
    HIPOK(wbfmm_hip_hel_m2l_mfma_f32(cls, cstart, ccount, td->d_start[level], td->d_end[level],
                                     level, N, (const float *)t->mps[level],
                                     (float *)t->mpr[level], od->m2l_plan_dev[N], od->f_hA[N],
                                     od->f_hA2[level], od->f_hCS[N], od->m2l_plan[N]));
    wbfmm_hip_free(cls);
  } else {
    hel_m2l64(t, od, level, tl, nt);
  }
  if (level == t->depth) return;
  {
    const int p32 = !td->h64[level], c32 = !td->h64[level + 1];
    const guint Nc = t->order_r[level + 1];
    if (p32 && c32 && Nr == Nc) {
      int N = (int)Nc;
      ensure_hel_f32_order(od, N);
      ensure_hel_f32_shift_tiles(od, t->D, level + 1, N, td->hk);
      HIPOK(wbfmm_hip_hel_shift_mfma_f32(1, tl, nt, NULL, NULL,
                                         td->have_targets ? td->d_flag[level + 1] : NULL, N,
                                         (const float *)t->mpr[level], (float *)t->mpr[level + 1],
                                         od->m2l_plan_dev[N], od->f_hshA[N], od->f_hDn[level + 1],
                                         od->f_hCSdn[N], od->m2l_plan[N]));
      return;
    }
    {
      double *src = p32 ? hel_level_to64(t, level, 1, tl, nt) : (double *)t->mpr[level];
      double *dst = (double *)t->mpr[level + 1];
      if (c32) {
        size_t bytes = ((size_t)1 << (3 * (level + 1))) * 2 * (Nc + 1) * (Nc + 1) * sizeof(double);
        HIPOK(wbfmm_hip_malloc((void **)&dst, bytes));
        HIPOK(wbfmm_hip_memset(dst, 0, bytes));
      }
      hel_l2l64(t, od, level, tl, nt, src, dst);
      if (c32) {
        uint32_t *cl;
        unsigned nc;
        if (td->have_targets) {
          cl = td->d_act[level + 1];
          nc = td->nact[level + 1];
        } else {
          cl = all_boxes(level + 1, &nc);
        }
        hel_level_to32(t, level + 1, 1, cl, nc, dst, 1);
        HIPOK(wbfmm_hip_free(dst));
      }
      if (p32) HIPOK(wbfmm_hip_free(src));
    }
  }
}

/* src/pass.c:448-537 (and the AVX variant, src/pass-avx.c): M2L at level,
   then L2L into level+1; work and nthreads are not used */
gint wbfmm_downward_pass_ref(wbfmm_tree_t *t, wbfmm_shift_operators_t *op,
                             guint level, gdouble *work, gint nthreads)
{
  (void)work;
  (void)nthreads;
  return hel_downward(t, op, level);
}

gint wbfmm_downward_pass_avx(wbfmm_tree_t *t, wbfmm_shift_operators_t *op,
                             guint level, gdouble *work, gint nthreads)
{
  (void)work;
  (void)nthreads;
  return hel_downward(t, op, level);
}

/* the field kernels: one thread per target (FP64 or mixed), or one
   wavefront per target box (WBFMM_HEL_FIELD=box, kept for comparison:
   slower on leaves of a few tens of targets) */
static int hel_field_box64(void)
{
  /* the calling thread's variant (wbfmm_hip_set_variant) first */
  const int v = wbfmm_hip_get_variant(WBFMM_VARIANT_FIELD);
  const char *e;
  if (v >= 0) return v == 1;
  e = getenv("WBFMM_HEL_FIELD");
  return e && strcmp(e, "box") == 0;
}

/* field values of n staged targets (Helmholtz), sorted by their Morton
   codes dcodes: fl bit 1 potential (2 doubles), bit 2 gradient (6),
   potential first */
static void hel_field_run(tree_dev *td, double k, const double *dxt,
                          const uint32_t *dtid, const uint64_t *dcodes, unsigned n,
                          double *src, gint sstr, double *d, gint dstr, int fl, double *f,
                          gint fstr, size_t nf_rows)
{
  wbfmm_tree_t *t = &td->t;
  unsigned lv = t->depth;
  double *q, *dp, *df;
  int staged, f32;
  size_t nelem;
  if (!t->mpr[lv]) WBFMM_FATAL("coefficients not initialised at level %u", lv);
  if (t->normals == NULL && d != NULL)
    WBFMM_FATAL("no normals in tree but dipole strengths specified (d != NULL)");
  f32 = level_f32(td, lv);
  if (f32 && (!td->hk_set || td->hk != k))
    WBFMM_FATAL("tree formed for k = %g, field asked for k = %g", td->hk, k);
  q = gather_strengths_n(td, src, sstr, 0, 2);
  dp = gather_strengths_n(td, d, dstr, 1, 2);
  nelem = nf_rows ? (nf_rows - 1) * (size_t)fstr + ((fl & 1) ? 2 : 0) + ((fl & 2) ? 6 : 0) : 0;
  df = field_begin(f, nelem, &staged);
  if (f32 && !hel_field_box64()) {
    /* mixed precision, one thread per target */
    HIPOK(wbfmm_hip_hel_field_f32(dxt, dtid, n, td->d_start[lv], td->d_end[lv], td->d_xyz, q,
                                  dp ? td->d_nrm : NULL, dp, lv, ORIGIN(t), t->D,
                                  t->order_r[lv], k, (const float *)t->mpr[lv], td->d_hsc[lv],
                                  fl, df, fstr));
  } else if (f32 || hel_field_box64()) {
    /* one wavefront per box holding targets (WBFMM_HEL_FIELD=box) */
    const size_t nb = (size_t)1 << (3 * lv);
    uint32_t *ts = NULL, *te = NULL, *tb = NULL;
    unsigned ntb = 0;
    HIPOK(wbfmm_hip_malloc((void **)&ts, sizeof(uint32_t) * nb));
    HIPOK(wbfmm_hip_malloc((void **)&te, sizeof(uint32_t) * nb));
    HIPOK(wbfmm_hip_malloc((void **)&tb, sizeof(uint32_t) * nb));
    HIPOK(wbfmm_hip_box_ranges(dcodes, n, lv, ts, te));
    HIPOK(wbfmm_hip_compact_boxes(ts, te, NULL, lv, tb, &ntb));
    HIPOK(wbfmm_hip_hel_field_box(tb, ntb, ts, te, dxt, dtid, td->d_start[lv], td->d_end[lv],
                                  td->d_xyz, q, dp ? td->d_nrm : NULL, dp, lv, ORIGIN(t), t->D,
                                  t->order_r[lv], k, f32 ? 32 : 64, t->mpr[lv],
                                  f32 ? td->d_hsc[lv] : NULL, fl, df, fstr));
    wbfmm_hip_free(ts);
    wbfmm_hip_free(te);
    wbfmm_hip_free(tb);
  } else {
    HIPOK(wbfmm_hip_hel_field(dxt, dtid, n, td->d_start[lv], td->d_end[lv], td->d_xyz, q,
                              dp ? td->d_nrm : NULL, dp, lv, ORIGIN(t), t->D,
                              t->order_r[lv], k, (double *)t->mpr[lv], fl, df, fstr));
  }
  field_finish(f, df, nelem, staged);
}

/* src/local.c:264-310: field at one point from the leaf box's local
   expansion and (eval_neighbours) the 27 neighbour boxes' sources;
   POTENTIAL: f[0..1], GRADIENT: f[0..5] (POTENTIAL|GRADIENT is refused as
   in the reference) */
gint wbfmm_tree_box_local_field(wbfmm_tree_t *t, guint level, guint b, gdouble k,
                                gdouble *x, gdouble *f, gint fstr, gdouble *src,
                                gint sstr, gdouble *d, gint dstr,
                                gboolean eval_neighbours, guint field,
                                gdouble *work)
{
  tree_dev *td = TD(t);
  double *dx, tmp[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  uint32_t *did;
  uint64_t *dc;
  int fl, i;
  (void)work;
  (void)b;
  if (t->problem != WBFMM_PROBLEM_HELMHOLTZ) WBFMM_FATAL("tree is not a Helmholtz tree");
  if (level != t->depth) WBFMM_FATAL("field evaluation is at the leaf level only");
  switch (field) {
    case WBFMM_FIELD_POTENTIAL:
      if (fstr < 2) WBFMM_FATAL("field stride (%d) too small for field calculation", fstr);
      fl = 1;
      break;
    case WBFMM_FIELD_GRADIENT:
      if (fstr < 6) WBFMM_FATAL("field stride (%d) too small for gradient calculation", fstr);
      fl = 2;
      break;
    default:
      WBFMM_FATAL("unrecognized field identifier (%u)", field);
  }
  stage_targets(x, 3 * sizeof(double), 1, &dx, &did, &dc, ORIGIN(t),
                t->D, 0);
  if (!eval_neighbours) { src = NULL; d = NULL; }
  hel_field_run(td, k, dx, did, dc, 1, src, sstr, d, dstr, fl, tmp, 8, 1);
  for (i = 0; i < (fl == 1 ? 2 : 6); i++) f[i] += tmp[i];
  wbfmm_hip_free(dx);
  wbfmm_hip_free(did);
  wbfmm_hip_free(dc);
  return 0;
}

/* addition: potential and/or gradient at ntgt points in one launch;
   f gets 2 (potential), 6 (gradient) or 8 (both, potential first)
   doubles per target at element stride fstr, accumulated */
gint wbfmm_helmholtz_field_targets(wbfmm_tree_t *t, gdouble k, gdouble *xt,
                                   gsize tstr, guint ntgt, gdouble *src, gint sstr,
                                   gdouble *d, gint dstr, guint field, gdouble *f,
                                   gint fstr)
{
  tree_dev *td = TD(t);
  double *dx;
  uint32_t *did;
  uint64_t *dc;
  int fl = field & 3;
  if (t->problem != WBFMM_PROBLEM_HELMHOLTZ) return -2;
  if (fl == 0 || (field & ~3u)) return -2;
  if (fstr < ((fl & 1) ? 2 : 0) + ((fl & 2) ? 6 : 0)) return -2;
  if (ntgt == 0) return 0;
  stage_targets(xt, tstr, ntgt, &dx, &did, &dc, ORIGIN(t), t->D, 1);
  hel_field_run(td, k, dx, did, dc, ntgt, src, sstr, d, dstr, fl, f, fstr, ntgt);
  HIPOK(wbfmm_hip_sync());
  wbfmm_hip_free(dx);
  wbfmm_hip_free(did);
  wbfmm_hip_free(dc);
  return 0;
}

/* src/direct.c:35-294: FP64 direct sum on the GPU at one field point;
   grad: gradient (6 doubles) instead of potential (2) */
static gint hel_total(gdouble k, gdouble *xs, gint xstride, gdouble *src, gint sstride,
                      gdouble *normals, gint nstr, gdouble *dipoles, gint dstr, gint nq,
                      gint nsrc, gdouble *xf, gdouble *field, int grad)
{
  double *dxs, *dq = NULL, *dd = NULL, *dn = NULL, *dxt, *df, tmp[8];
  uint32_t *ids;
  uint32_t *h;
  gint i;
  if (nq != 1) WBFMM_FATAL("only nq == 1 is supported");
  if (nsrc <= 0) return 0;
  HIPOK(wbfmm_hip_malloc((void **)&dxs, sizeof(double) * 4 * nsrc));
  HIPOK(wbfmm_hip_stage4(xs, xstride * sizeof(double), nsrc, dxs));
  HIPOK(wbfmm_hip_malloc((void **)&ids, sizeof(uint32_t) * nsrc));
  h = malloc(sizeof(uint32_t) * nsrc);
  for (i = 0; i < nsrc; i++) h[i] = i;
  HIPOK(wbfmm_hip_memcpy(ids, h, sizeof(uint32_t) * nsrc));
  free(h);
  if (src) {
    HIPOK(wbfmm_hip_malloc((void **)&dq, sizeof(double) * 2 * nsrc));
    HIPOK(wbfmm_hip_gather_strided(src, sstride, 2, ids, nsrc, dq));
  }
  if (dipoles) {
    if (!normals) WBFMM_FATAL("dipole strengths without normals");
    HIPOK(wbfmm_hip_malloc((void **)&dd, sizeof(double) * 2 * nsrc));
    HIPOK(wbfmm_hip_gather_strided(dipoles, dstr, 2, ids, nsrc, dd));
    HIPOK(wbfmm_hip_malloc((void **)&dn, sizeof(double) * 4 * nsrc));
    HIPOK(wbfmm_hip_stage4(normals, nstr * sizeof(double), nsrc, dn));
  }
  HIPOK(wbfmm_hip_malloc((void **)&dxt, sizeof(double) * 4));
  {
    double p4[4] = {xf[0], xf[1], xf[2], 0.0};
    HIPOK(wbfmm_hip_memcpy(dxt, p4, sizeof(p4)));
  }
  HIPOK(wbfmm_hip_malloc((void **)&df, sizeof(double) * 8));
  HIPOK(wbfmm_hip_memset(df, 0, sizeof(double) * 8));
  HIPOK(wbfmm_hip_hel_direct(dxs, dq, dn, dd, (unsigned)nsrc, dxt, 1, k, grad ? 2 : 1, 0.0,
                             df, 8));
  HIPOK(wbfmm_hip_memcpy(tmp, df, sizeof(double) * 8));
  for (i = 0; i < (grad ? 6 : 2); i++) field[i] += tmp[i];
  wbfmm_hip_free(dxs); wbfmm_hip_free(ids); wbfmm_hip_free(dq);
  wbfmm_hip_free(dd); wbfmm_hip_free(dn); wbfmm_hip_free(dxt); wbfmm_hip_free(df);
  return 0;
}

gint wbfmm_total_field(gdouble k, gdouble *xs, gint xstride, gdouble *src,
                       gint sstride, gdouble *normals, gint nstr, gdouble *dipoles,
                       gint dstr, gint nq, gint nsrc, gdouble *xf, gdouble *field,
                       gint fstr)
{
  (void)fstr;
  return hel_total(k, xs, xstride, src, sstride, normals, nstr, dipoles, dstr, nq, nsrc,
                   xf, field, 0);
}

gint wbfmm_total_field_grad(gdouble k, gdouble *xs, gint xstride, gdouble *src,
                            gint sstride, gdouble *normals, gint nstr, gdouble *dipoles,
                            gint dstr, gint nq, gint nsrc, gdouble *xf, gdouble *field,
                            gint fstr)
{
  (void)fstr;
  return hel_total(k, xs, xstride, src, sstride, normals, nstr, dipoles, dstr, nq, nsrc,
                   xf, field, 1);
}

/* ------------------------------------------------------------------ */
/* Single-precision entry points of the Laplace tree pipeline, the      */
/* target lists and the direct sum (src/wbfmm.h:432-440, :690-740,      */
/* :996-1112).  A tree made by wbfmm_tree_new_f is an FP32 tree (as     */
/* wbfmm_tree_set_precision(t, 32)): expansions in FP32 scaled by the   */
/* box width and FP32 matrix-core translations, orders 2..15 and the    */
/* same order at every level, with FP64 recurrences and near field.     */
/* Float arguments are host arrays, widened to double on the host;      */
/* float outputs are accumulated in float.  The public origin holds     */
/* gfloat like the reference's; the passes use a double copy.           */
/* ------------------------------------------------------------------ */

/* n items of ncomp floats, item i at x + i * stride floats, into n *
   ncomp contiguous doubles (NULL for NULL) */
static double *widen(const float *x, size_t stride, int ncomp, size_t n)
{
  double *y;
  size_t i;
  int c;
  if (!x) return NULL;
  if (wbfmm_hip_is_device(x)) WBFMM_FATAL("the single-precision entry points take host arrays");
  y = malloc(sizeof(double) * ncomp * (n ? n : 1));
  if (!y) WBFMM_FATAL("out of host memory");
  for (i = 0; i < n; i++)
    for (c = 0; c < ncomp; c++) y[i * ncomp + c] = x[i * stride + c];
  return y;
}

static void require_f(wbfmm_tree_t *t)
{
  if (t->size != sizeof(gfloat)) WBFMM_FATAL("not a single-precision tree (wbfmm_tree_new_f)");
}

wbfmm_tree_t *wbfmm_tree_new_f(gfloat *x, gfloat D, guint maxpoints)
{
  double x0[3] = {x[0], x[1], x[2]};
  wbfmm_tree_t *t = wbfmm_tree_new(x0, (double)D, maxpoints);
  gfloat *xf = (gfloat *)wbfmm_tree_origin(t);
  TD(t)->prec32 = 1;
  TD(t)->host_mirror = 0;
  t->size = sizeof(gfloat);
  /* the public origin in the tree's precision (src/boxes.c:58-88) */
  memset(t->x, 0, sizeof(t->x));
  xf[0] = x[0];
  xf[1] = x[1];
  xf[2] = x[2];
  return t;
}

gint wbfmm_tree_add_points_f(wbfmm_tree_t *t, gpointer pts, gsize pstr, gpointer normals,
                             gsize nstr, guint npts, gboolean sorted)
{
  double *xd, *nd;
  require_f(t);
  if (pstr % sizeof(gfloat) || nstr % sizeof(gfloat))
    WBFMM_FATAL("point strides must be whole floats");
  xd = widen((const float *)pts, pstr / sizeof(gfloat), 3, npts);
  nd = widen((const float *)normals, nstr / sizeof(gfloat), 3, npts);
  add_points_dbl(t, xd, 3 * sizeof(double), nd, 3 * sizeof(double), npts, sorted);
  HIPOK(wbfmm_hip_sync());
  free(xd);
  free(nd);
  t->points = (char *)pts;
  t->pstr = pstr;
  t->normals = (char *)normals;
  t->nstr = nstr;
  return 0;
}

gint wbfmm_tree_sort_points_f(wbfmm_tree_t *t, gpointer pts, gsize psize, guint npts)
{
  double *xd;
  gint rc;
  require_f(t);
  if (psize % sizeof(gfloat)) WBFMM_FATAL("point size must be whole floats");
  /* the Morton codes from the coordinates widened to double, as in
     wbfmm_tree_add_points_f */
  xd = widen((const float *)pts, psize / sizeof(gfloat), 3, npts);
  rc = sort_records(t, xd, 3 * sizeof(double), pts, psize, npts);
  free(xd);
  return rc;
}

gint wbfmm_tree_refine_f(wbfmm_tree_t *t)
{
  require_f(t);
  return wbfmm_tree_refine(t);
}

gint wbfmm_tree_laplace_coefficient_init_f(wbfmm_tree_t *t, guint l, guint nr, guint ns)
{
  require_f(t);
  return wbfmm_tree_laplace_coefficient_init(t, l, nr, ns);
}

gint wbfmm_tree_laplace_leaf_expansions_f(wbfmm_tree_t *t, gfloat *src, gint sstr,
                                          gfloat *dipoles, gint dstr,
                                          gboolean zero_expansions, gfloat *work)
{
  double *q, *d;
  (void)work;
  require_f(t);
  q = widen(src, sstr, 1, t->npoints);
  d = widen(dipoles, dstr, 1, t->npoints);
  wbfmm_tree_laplace_leaf_expansions(t, q, 1, d, 1, zero_expansions, NULL);
  HIPOK(wbfmm_hip_sync());
  free(q);
  free(d);
  return 0;
}

gint wbfmm_shift_angle_table_init_f(void) { return 0; }

wbfmm_shift_operators_t *wbfmm_shift_operators_new_f(guint L, gboolean bw, gfloat *work)
{
  (void)work;
  return wbfmm_shift_operators_new(L, bw, NULL);
}

gint wbfmm_laplace_coaxial_translate_init_f(gint N)
{
  return wbfmm_laplace_coaxial_translate_init(N);
}

gint wbfmm_laplace_upward_pass_f(wbfmm_tree_t *t, wbfmm_shift_operators_t *op, guint level,
                                 gfloat *work)
{
  (void)work;
  require_f(t);
  return wbfmm_laplace_upward_pass(t, op, level, NULL);
}

gint wbfmm_laplace_downward_pass_ref_f(wbfmm_tree_t *t, wbfmm_shift_operators_t *op,
                                       guint level, gfloat *work, gint nthreads)
{
  (void)work;
  require_f(t);
  return wbfmm_laplace_downward_pass_ref(t, op, level, NULL, nthreads);
}

/* values box_field writes at unit stride: 1 (potential) or 3 (gradient) */
static int lap_field_count(guint field)
{
  return (field & WBFMM_FIELD_GRADIENT) ? 3 : 1;
}

gint wbfmm_laplace_box_field_f(wbfmm_tree_t *t, guint level, guint b, gfloat *src, gint sstr,
                               gfloat *d, gint dstr, guint field, gboolean eval_neighbours,
                               gfloat *x, gfloat *f, gint fstr, gfloat *work)
{
  double *q, *dd, xd[3] = {x[0], x[1], x[2]}, tmp[4] = {0, 0, 0, 0};
  int k;
  (void)work;
  require_f(t);
  q = eval_neighbours ? widen(src, sstr, 1, t->npoints) : NULL;
  dd = eval_neighbours ? widen(d, dstr, 1, t->npoints) : NULL;
  wbfmm_laplace_box_field(t, level, b, q, 1, dd, 1, field, eval_neighbours, xd, tmp, fstr,
                          NULL);
  HIPOK(wbfmm_hip_sync());
  free(q);
  free(dd);
  for (k = 0; k < lap_field_count(field); k++) f[k] += (gfloat)tmp[k];
  return 0;
}

gint wbfmm_laplace_field_direct_f(gfloat *xs, gint xstride, gfloat *n, gint nstr, gint nsrc,
                                  gfloat *src, gint sstr, gfloat *d, gint dstr, gint nq,
                                  guint field, gfloat *xf, gfloat *f, gint fstr)
{
  double *xd, *nd, *q, *dd, xfd[3] = {xf[0], xf[1], xf[2]}, tmp[4] = {0, 0, 0, 0};
  int k;
  xd = widen(xs, xstride, 3, nsrc);
  nd = widen(n, nstr, 3, nsrc);
  q = widen(src, sstr, 1, nsrc);
  dd = widen(d, dstr, 1, nsrc);
  wbfmm_laplace_field_direct(xd, 3, nd, 3, nsrc, q, 1, dd, 1, nq, field, xfd, tmp, fstr);
  HIPOK(wbfmm_hip_sync());
  free(xd);
  free(nd);
  free(q);
  free(dd);
  for (k = 0; k < lap_field_count(field); k++) f[k] += (gfloat)tmp[k];
  return 0;
}

wbfmm_target_list_t *wbfmm_target_list_new_f(wbfmm_tree_t *t, guint npts)
{
  wbfmm_target_list_t *l;
  require_f(t);
  l = wbfmm_target_list_new(t, npts);
  l->size = sizeof(gfloat);
  return l;
}

gint wbfmm_target_list_add_points_f(wbfmm_target_list_t *l, gpointer pts, gsize pstr,
                                    guint npts)
{
  double *xd;
  if (pstr % sizeof(gfloat)) WBFMM_FATAL("point stride must be whole floats");
  xd = widen((const float *)pts, pstr / sizeof(gfloat), 3, npts);
  wbfmm_target_list_add_points(l, xd, 3 * sizeof(double), npts);
  HIPOK(wbfmm_hip_sync());
  free(xd);
  l->points = (char *)pts;
  l->pstr = pstr;
  return 0;
}

gint wbfmm_laplace_target_list_local_coefficients_f(wbfmm_target_list_t *l, guint source,
                                                    gfloat *work)
{
  (void)work;
  return wbfmm_laplace_target_list_local_coefficients(l, source, NULL);
}

gint wbfmm_target_list_local_field_f(wbfmm_target_list_t *l, gfloat *src, gint sstr,
                                     gfloat *d, gint dstr, gfloat *f, gint fstr)
{
  size_t nt = l->npoints, i;
  int fl = (int)(l->field & 3), nv, k;
  /* Laplace: real strengths, 1 + 3 values; Helmholtz: complex, 2 + 6 */
  const int cw = l->t->problem == WBFMM_PROBLEM_HELMHOLTZ ? 2 : 1;
  double *q, *dd, *fd;
  nv = ((fl & 1) ? cw : 0) + ((fl & 2) ? 3 * cw : 0);
  if (fstr < nv) WBFMM_FATAL("field stride %d too small", fstr);
  q = widen(src, sstr, cw, l->t->npoints);
  dd = widen(d, dstr, cw, l->t->npoints);
  fd = calloc((nt ? nt : 1) * nv, sizeof(double));
  if (!fd) WBFMM_FATAL("out of host memory");
  wbfmm_target_list_local_field(l, q, cw, dd, cw, fd, nv);
  HIPOK(wbfmm_hip_sync());
  for (i = 0; i < nt; i++)
    for (k = 0; k < nv; k++) f[i * fstr + k] += (gfloat)fd[i * nv + k];
  free(q);
  free(dd);
  free(fd);
  return 0;
}

/* ------------------------------------------------------------------ */
/* Single-precision entry points of the Helmholtz pipeline              */
/* (src/wbfmm.h:869-1112).  A tree made by wbfmm_tree_new_f runs the    */
/* mixed-precision Helmholtz path: FP32 expansions scaled per degree    */
/* and FP32 matrix-core translations at the levels where k a_l is small */
/* and the orders are served, FP64 elsewhere, with FP64 tables and      */
/* recurrences.  Float arguments are host arrays widened to double on   */
/* the host (complex strengths: 2 floats at element stride); float      */
/* outputs are accumulated in float.                                    */
/* ------------------------------------------------------------------ */

gint wbfmm_tree_coefficient_init_f(wbfmm_tree_t *t, guint l, guint nr, guint ns)
{
  require_f(t);
  return wbfmm_tree_coefficient_init(t, l, nr, ns);
}

gint wbfmm_shift_operators_coaxial_SR_init_f(wbfmm_shift_operators_t *w, gfloat D, guint level,
                                             guint L, gfloat k, gfloat *work)
{
  (void)work;
  return wbfmm_shift_operators_coaxial_SR_init(w, (double)D, level, L, (double)k, NULL);
}

gint wbfmm_shift_operators_coaxial_SS_init_f(wbfmm_shift_operators_t *w, gfloat D, guint level,
                                             guint L, gfloat k, gfloat *work)
{
  (void)work;
  return wbfmm_shift_operators_coaxial_SS_init(w, (double)D, level, L, (double)k, NULL);
}

gint wbfmm_tree_leaf_expansions_f(wbfmm_tree_t *t, gfloat k, gfloat *src, gint sstr,
                                  gfloat *dipoles, gint dstr, gboolean zero_expansions,
                                  gfloat *work)
{
  double *q, *d;
  (void)work;
  require_f(t);
  q = widen(src, sstr, 2, t->npoints);
  d = widen(dipoles, dstr, 2, t->npoints);
  wbfmm_tree_leaf_expansions(t, (double)k, q, 2, d, 2, zero_expansions, NULL);
  HIPOK(wbfmm_hip_sync());
  free(q);
  free(d);
  return 0;
}

gint wbfmm_upward_pass_f(wbfmm_tree_t *t, wbfmm_shift_operators_t *op, guint level,
                         gfloat *work)
{
  (void)work;
  require_f(t);
  return wbfmm_upward_pass(t, op, level, NULL);
}

gint wbfmm_downward_pass_ref_f(wbfmm_tree_t *t, wbfmm_shift_operators_t *op, guint level,
                               gfloat *work, gint nthreads)
{
  (void)work;
  require_f(t);
  return wbfmm_downward_pass_ref(t, op, level, NULL, nthreads);
}

gint wbfmm_tree_box_local_field_f(wbfmm_tree_t *t, guint level, guint b, gfloat k, gfloat *x,
                                  gfloat *f, gint fstr, gfloat *src, gint sstr, gfloat *d,
                                  gint dstr, gboolean eval_neighbours, guint field,
                                  gfloat *work)
{
  double *q, *dd, xd[3] = {x[0], x[1], x[2]}, tmp[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  const int nv = (field & WBFMM_FIELD_GRADIENT) ? 6 : 2;
  int i;
  (void)work;
  require_f(t);
  q = eval_neighbours ? widen(src, sstr, 2, t->npoints) : NULL;
  dd = eval_neighbours ? widen(d, dstr, 2, t->npoints) : NULL;
  wbfmm_tree_box_local_field(t, level, b, (double)k, xd, tmp, fstr, q, 2, dd, 2,
                             eval_neighbours, field, NULL);
  HIPOK(wbfmm_hip_sync());
  free(q);
  free(dd);
  for (i = 0; i < nv; i++) f[i] += (gfloat)tmp[i];
  return 0;
}

gint wbfmm_target_list_local_coefficients_f(wbfmm_target_list_t *l, gfloat k, gfloat *work)
{
  (void)work;
  return wbfmm_target_list_local_coefficients(l, (double)k, NULL);
}

/* the direct sums at one point (src/direct.c:35-294): 2 (potential) or 6
   (gradient) floats */
static gint hel_total_f(gfloat k, gfloat *xs, gint xstride, gfloat *src, gint sstride,
                        gfloat *normals, gint nstr, gfloat *dipoles, gint dstr, gint nq,
                        gint nsrc, gfloat *xf, gfloat *field, int grad)
{
  double *xd, *nd, *q, *dd, xfd[3] = {xf[0], xf[1], xf[2]}, tmp[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  int i;
  if (nq != 1) WBFMM_FATAL("only nq == 1 is supported");
  xd = widen(xs, xstride, 3, nsrc);
  nd = widen(normals, nstr, 3, nsrc);
  q = widen(src, sstride, 2, nsrc);
  dd = widen(dipoles, dstr, 2, nsrc);
  hel_total((double)k, xd, 3, q, 2, nd, 3, dd, 2, 1, nsrc, xfd, tmp, grad);
  HIPOK(wbfmm_hip_sync());
  free(xd);
  free(nd);
  free(q);
  free(dd);
  for (i = 0; i < (grad ? 6 : 2); i++) field[i] += (gfloat)tmp[i];
  return 0;
}

gint wbfmm_total_field_f(gfloat k, gfloat *xs, gint xstride, gfloat *src, gint sstride,
                         gfloat *normals, gint nstr, gfloat *dipoles, gint dstr, gint nq,
                         gint nsrc, gfloat *xf, gfloat *field, gint fstr)
{
  (void)fstr;
  return hel_total_f(k, xs, xstride, src, sstride, normals, nstr, dipoles, dstr, nq, nsrc, xf,
                     field, 0);
}

gint wbfmm_total_field_grad_f(gfloat k, gfloat *xs, gint xstride, gfloat *src, gint sstride,
                              gfloat *normals, gint nstr, gfloat *dipoles, gint dstr, gint nq,
                              gint nsrc, gfloat *xf, gfloat *field, gint fstr)
{
  (void)fstr;
  return hel_total_f(k, xs, xstride, src, sstride, normals, nstr, dipoles, dstr, nq, nsrc, xf,
                     field, 1);
}
