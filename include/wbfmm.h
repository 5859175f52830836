/*
 * wbfmm.h -- drop-in C API of the MI355X (gfx950) implementation of the
 * wbfmm octree FMM evaluation path.
 *
 * Callers of the reference library (mjcarley/wbfmm, src/wbfmm.h) include
 * this header instead and link -lwbfmm.  The struct layouts and the
 * signatures below are those of the reference header; every declaration
 * cites the reference declaration it replaces.  GLib is not required: the
 * GLib scalar type names used by the reference are defined here as plain C
 * types of the same size (unless glib.h was already included).
 *
 * Differences in behaviour are listed next to each function; the main ones:
 *   - expansion coefficients (mps/mpr, the box pointers into them) live in
 *     GPU memory (HBM).  They keep the reference's packing and
 *     normalisation; read them with wbfmm_tree_coefficients_download();
 *   - point, normal, strength and field arrays may be host or device
 *     memory; device arrays are used in place, host arrays are copied;
 *   - the number of source components nq must be 1;
 *   - errors that the reference turns into g_error() aborts print a message
 *     to stderr and abort as well; entry points added by this library
 *     (wbfmm_hip_*, *_targets) return a negative status instead.
 */

#ifndef WBFMM_H_INCLUDED
#define WBFMM_H_INCLUDED

#include <stdio.h>
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#ifndef __G_LIB_H__
typedef int gint;
typedef unsigned int guint;
typedef int gboolean;
typedef double gdouble;
typedef float gfloat;
typedef void *gpointer;
typedef size_t gsize;
typedef uint32_t guint32;
typedef uint64_t guint64;
typedef int64_t gint64;
#ifndef TRUE
#define TRUE 1
#endif
#ifndef FALSE
#define FALSE 0
#endif
#endif

/* src/wbfmm.h:34-39 */
#define WBFMM_INDEX_SCALE (1LU << 63)
#define WBFMM_INDEX_SHIFT (1U << 20)
#define WBFMM_LOCAL_CUTOFF_RADIUS 1e-6
#ifndef WBFMM_THREAD_NUMBER_MAX
#define WBFMM_THREAD_NUMBER_MAX 16
#endif

/* src/wbfmm.h:69-74 */
typedef struct {
  guint32 i, n;
  gpointer mps, mpr;
} wbfmm_box_t;

/* src/wbfmm.h:77 */
#define WBFMM_TREE_MAX_DEPTH 16

/* src/wbfmm.h:86-90 */
typedef enum {
  WBFMM_PROBLEM_LAPLACE = 1,
  WBFMM_PROBLEM_HELMHOLTZ = 2
} wbfmm_problem_t;

/* src/wbfmm.h:99-106 */
typedef enum {
  WBFMM_FIELD_POTENTIAL = 1 << 0,
  WBFMM_FIELD_GRADIENT = 1 << 1,
  WBFMM_FIELD_CURL = 1 << 2,
  WBFMM_FIELD_CURL_GRADIENT = 1 << 3
} wbfmm_field_t;

/* src/wbfmm.h:116-145 (layout kept field for field) */
typedef struct {
  wbfmm_problem_t problem;
  wbfmm_box_t *boxes[WBFMM_TREE_MAX_DEPTH + 1];
  gboolean sorted;
  guint maxpoints, npoints, *ip, nq, depth,
      order_s[WBFMM_TREE_MAX_DEPTH + 1], order_r[WBFMM_TREE_MAX_DEPTH + 1];
  char x[24], *normals, *points;
  gpointer mps[WBFMM_TREE_MAX_DEPTH + 1], mpr[WBFMM_TREE_MAX_DEPTH + 1];
  gsize size, nstr, pstr;
  gdouble D;
} wbfmm_tree_t;

/* src/wbfmm.h:147-151 */
typedef enum {
  WBFMM_SOURCE_MONOPOLE = 1 << 0,
  WBFMM_SOURCE_DIPOLE = 1 << 1
} wbfmm_source_t;

/* src/wbfmm.h:161-188 (layout kept) */
typedef struct {
  wbfmm_tree_t *t;
  guint field, source, maxpoints, npoints, *ip, nc;
  guint32 *boxes;
  char *points;
  gsize size, pstr;
  gint *ibox, *isrc, *ics;
  gpointer cfft, csrc;
  gboolean complex;
} wbfmm_target_list_t;

/* src/wbfmm.h:199-215 (layout kept) */
typedef struct {
  gsize size;
  guint Lmax, nlevels, L[WBFMM_TREE_MAX_DEPTH + 1], nerot;
  gpointer SR[WBFMM_TREE_MAX_DEPTH + 1], SS[WBFMM_TREE_MAX_DEPTH + 1],
      rotations;
  gboolean bw;
} wbfmm_shift_operators_t;

/* accessor macros, src/wbfmm.h:217-245 */
#define wbfmm_tree_point_number(_t) ((_t)->npoints)
#define wbfmm_tree_point_number_max(_t) ((_t)->maxpoints)
#define wbfmm_tree_depth(_t) ((_t)->depth)
#define wbfmm_tree_boxes_level(_t, _l) ((_t)->boxes[(_l)])
#define wbfmm_tree_width(_t) ((_t)->D)
#define wbfmm_tree_origin(_t) ((gpointer)(&((_t)->x[0])))
#define wbfmm_tree_problem(_t) ((_t)->problem)
#define wbfmm_tree_source_size(_t) ((_t)->nq)
#define wbfmm_target_list_tree(_t) ((_t)->t)
#define wbfmm_target_list_point_number(_t) ((_t)->npoints)
#define wbfmm_target_list_point_number_max(_t) ((_t)->maxpoints)
#define wbfmm_target_list_field(_l) ((_l)->field)
#define wbfmm_target_point_box(_l, _i) ((_l)->boxes[(_i)])

#define wbfmm_element_number_coaxial(_N) \
  (((_N) + 1) * ((_N) + 2) * ((_N) + 3) / 6 + ((_N) + 1) * ((_N) + 2) / 2 + (_N) + 1)
#define wbfmm_element_number_rotation(_N) \
  ((((_N) + 1) * (4 * (_N) + 3) * ((_N) + 2) / 6 + (((_N) + 1) + ((_N + 1))) * ((_N + 1) + 1) + (_N)))
#define wbfmm_T_rotation_matrix_size(_n) (((_n) + 1) * (2 * (_n) + 1) * (2 * (_n) + 3) / 3)
#define wbfmm_box_index_parent(_i) ((_i) / 8)
#define wbfmm_box_index_first_child(_i) ((_i) * 8)
#define wbfmm_coefficient_index_nm(_n, _m) ((_n) * ((_n) + 1) + (_m))
#define wbfmm_coefficient_number(_n) wbfmm_coefficient_index_nm(((_n) + 1), (-((_n) + 1)))
#define wbfmm_conjugate_index_nm(_n, _m) (((_n) * ((_n) + 1) / 2) + (_m))

/* ------------------------------------------------------------------ */
/* indexing, precision independent (src/wbfmm.h:1114-1135)           */
/* ------------------------------------------------------------------ */
guint64 wbfmm_point_locate_box(guint64 x, guint level);               /* src/wbfmm.h:1115 */
gint wbfmm_point_from_index(guint64 i, guint32 *x, guint32 *y, guint32 *z); /* :1116 */
guint64 wbfmm_box_index(guint32 i, guint32 j, guint32 k);             /* src/wbfmm.h:1117 */
gint wbfmm_box_location(guint64 idx, guint32 *i, guint32 *j, guint32 *k); /* :1118 */
guint64 wbfmm_box_parent(guint64 idx);                                /* src/wbfmm.h:1119 */
guint64 wbfmm_box_first_child(guint64 idx);                           /* src/wbfmm.h:1120 */
gint wbfmm_box_neighbours(guint level, guint64 idx, guint64 *neighbours); /* :1128 */
gint wbfmm_box_interaction_list_4(guint level, guint64 idx, guint64 *list,
                                  gboolean sort);                     /* src/wbfmm.h:1129-1130 */
gint wbfmm_box_interaction_index(gint i, gint j, gint k);             /* src/wbfmm.h:1132 */

/* ------------------------------------------------------------------ */
/* tree (double precision)                                            */
/* ------------------------------------------------------------------ */
/* src/wbfmm.h:835; the tree's device state is created with it */
wbfmm_tree_t *wbfmm_tree_new(gdouble *x, gdouble D, guint maxpoints);
/* src/wbfmm.h:810-813.  pts/normals: host or device, byte strides.  The
   pointers are kept (borrowed, as in the reference); a device copy of the
   coordinates is taken, sorted by Morton index on the GPU (stable).
   Adding points to a tree that already holds some starts it afresh
   (depth 0, target lists to be added again), so one tree and its device
   buffers can serve repeated solves. */
gint wbfmm_tree_add_points(wbfmm_tree_t *t, gpointer pts, gsize pstr,
                           gpointer normals, gsize nstr, guint npts,
                           gboolean sorted);
/* src/wbfmm.h:826-828 (src/boxes.c:190-206): reorders the caller's npts
   records of psize bytes (coordinates first) in place by Morton index,
   equal indices keeping their order; Morton codes and the stable sort run
   on the GPU, the records move on the GPU for device arrays and on the
   host for host arrays.  Aborts on a point outside the domain. */
gint wbfmm_tree_sort_points(wbfmm_tree_t *t, gpointer pts, gsize psize, guint npts);
/* src/wbfmm.h:808: adds one level; box ranges computed on the GPU */
gint wbfmm_tree_refine(wbfmm_tree_t *t);
/* src/wbfmm.h:807 */
guint64 wbfmm_point_box(wbfmm_tree_t *t, guint level, gdouble *x);
/* src/wbfmm.h:831 */
guint64 wbfmm_point_index_3d(gdouble *x, gdouble *c, gdouble D);
/* src/wbfmm.h:846-849 */
gint wbfmm_box_location_from_index(guint64 i, guint32 level, gdouble *x0,
                                   gdouble D, gdouble *x, gdouble *wb);
/* src/wbfmm.h:850-852 */
gint wbfmm_tree_box_centre(wbfmm_tree_t *t, guint level, guint64 b,
                           gdouble *xb, gdouble *wb);
/* src/wbfmm.h:853-857 */
gint wbfmm_points_origin_width(gdouble *x, gint str, gint n, gdouble *xmin,
                               gdouble *xmax, gdouble *D,
                               gboolean init_limits);
/* src/wbfmm.h:1134: zeroes mps and mpr at one level (device memset) */
gint wbfmm_tree_coefficients_zero(wbfmm_tree_t *t, guint level);
/* src/wbfmm.h:844-845 */
gint wbfmm_truncation_number(wbfmm_tree_t *t, gdouble k, guint level,
                             gdouble tol);

/* ------------------------------------------------------------------ */
/* shift operators                                                    */
/* ------------------------------------------------------------------ */
/* src/wbfmm.h:793 */
gint wbfmm_shift_angle_table_init(void);
/* src/wbfmm.h:794-796: rotation operators up to order L; device tables
   for the M2M/M2L/L2L kernels are built from them when a pass first
   needs a given order.  bw selects no different arithmetic here: the
   GPU passes are exact for either setting. */
wbfmm_shift_operators_t *wbfmm_shift_operators_new(guint L, gboolean bw,
                                                   gdouble *work);
/* src/wbfmm.h:791-792 */
gint wbfmm_shift_angles_list4(gint i, gint j, gint k, gdouble *th,
                              gdouble *ph, gdouble *ch, gdouble *rs);

/* ------------------------------------------------------------------ */
/* Laplace (double precision)                                         */
/* ------------------------------------------------------------------ */
/* src/wbfmm.h:580: Laplace coaxial tables up to order N (process global,
   as in the reference) */
gint wbfmm_laplace_coaxial_translate_init(gint N);
/* src/wbfmm.h:679-682: device coefficient arrays for level l */
gint wbfmm_tree_laplace_coefficient_init(wbfmm_tree_t *t, guint l, guint nr,
                                         guint ns);
/* src/wbfmm.h:687-693: P2M on the GPU.  src, dipoles: host or device */
gint wbfmm_tree_laplace_leaf_expansions(wbfmm_tree_t *t, gdouble *src,
                                        gint sstr, gdouble *dipoles,
                                        gint dstr, gboolean zero_expansions,
                                        gdouble *work);
/* src/wbfmm.h:713-716: M2M from level into level-1 */
gint wbfmm_laplace_upward_pass(wbfmm_tree_t *t, wbfmm_shift_operators_t *op,
                               guint level, gdouble *work);
/* src/wbfmm.h:701-708: M2L at level, then L2L to level+1; nthreads is
   ignored (the GPU runs every box).  _ref and _avx are the same GPU code. */
gint wbfmm_laplace_downward_pass_ref(wbfmm_tree_t *t,
                                     wbfmm_shift_operators_t *op,
                                     guint level, gdouble *work,
                                     gint nthreads);
gint wbfmm_laplace_downward_pass_avx(wbfmm_tree_t *t,
                                     wbfmm_shift_operators_t *op,
                                     guint level, gdouble *work,
                                     gint nthreads);
#define wbfmm_laplace_downward_pass(_t, _op, _level, _work, _nth) \
  wbfmm_laplace_downward_pass_ref(_t, _op, _level, _work, _nth)
/* src/wbfmm.h:723-728: field at one point (one small GPU launch; use the
   target-list or wbfmm_laplace_field_targets for many points).  As in the
   reference, field == POTENTIAL|GRADIENT returns the gradient only. */
gint wbfmm_laplace_box_field(wbfmm_tree_t *t, guint level, guint b,
                             gdouble *src, gint sstr, gdouble *d, gint dstr,
                             guint field, gboolean eval_neighbours,
                             gdouble *x, gdouble *f, gint fstr,
                             gdouble *work);
/* src/wbfmm.h:402-409: FP64 direct sum on the GPU (no self cutoff, as in
   the reference; the dipole term carries the 1/4pi factor) */
gint wbfmm_laplace_field_direct(gdouble *xs, gint xstride, gdouble *n,
                                gint nstr, gint nsrc, gdouble *src,
                                gint sstr, gdouble *d, gint dstr, gint nq,
                                guint field, gdouble *xf, gdouble *f,
                                gint fstr);

/* ------------------------------------------------------------------ */
/* target lists (src/wbfmm.h, src/tree.c:86-122, src/targets.c)       */
/* ------------------------------------------------------------------ */
/* ---- Helmholtz path (complex coefficients, FP64) ----------------------- */
/* src/wbfmm.h:836-838: complex coefficient arrays of level l (device) */
gint wbfmm_tree_coefficient_init(wbfmm_tree_t *t, guint l, guint nr, guint ns);
/* src/wbfmm.h:839-843: P2M on the GPU; src, dipoles: complex strengths
   (2 doubles) at element stride, host or device; dipoles need normals */
gint wbfmm_tree_leaf_expansions(wbfmm_tree_t *t, gdouble k, gdouble *src,
                                gint sstr, gdouble *dipoles, gint dstr,
                                gboolean zero_expansions, gdouble *work);
/* src/wbfmm.h:792-797: coaxial tables of a level, built on the host in
   quad precision and kept on the device (w->SR[level], w->SS[level] are
   device pointers); one wavenumber per operator set */
gint wbfmm_shift_operators_coaxial_SR_init(wbfmm_shift_operators_t *w,
                                           gdouble D, guint level, guint L,
                                           gdouble k, gdouble *work);
gint wbfmm_shift_operators_coaxial_SS_init(wbfmm_shift_operators_t *w,
                                           gdouble D, guint level, guint L,
                                           gdouble k, gdouble *work);
/* src/wbfmm.h:798-806: M2M from level into level-1; M2L at level then L2L
   into level+1 (nthreads and work are not used) */
gint wbfmm_upward_pass(wbfmm_tree_t *t, wbfmm_shift_operators_t *op,
                       guint level, gdouble *work);
gint wbfmm_downward_pass_ref(wbfmm_tree_t *t, wbfmm_shift_operators_t *op,
                             guint level, gdouble *work, gint nthreads);
gint wbfmm_downward_pass_avx(wbfmm_tree_t *t, wbfmm_shift_operators_t *op,
                             guint level, gdouble *work, gint nthreads);
#define wbfmm_downward_pass(_t, _op, _level, _work, _nth) \
  wbfmm_downward_pass_ref(_t, _op, _level, _work, _nth)
/* src/wbfmm.h:810-818: field at one point of leaf box b; POTENTIAL:
   f[0..1], GRADIENT: f[0..5]; POTENTIAL|GRADIENT is refused as in the
   reference (src/local.c:297-300) */
gint wbfmm_tree_box_local_field(wbfmm_tree_t *t, guint level, guint b,
                                gdouble k, gdouble *x, gdouble *f, gint fstr,
                                gdouble *src, gint sstr, gdouble *d,
                                gint dstr, gboolean eval_neighbours,
                                guint field, gdouble *work);
/* src/wbfmm.h:317-330: FP64 direct sum on the GPU at one point (potential
   f[0..1], or gradient f[0..5] for _grad), G = h0(kr)/(4 pi), dipole term
   d nrm.grad_x G; values are added to field (the reference's mixed path
   divides the whole field array by 4 pi instead, src/direct.c:134) */
gint wbfmm_total_field(gdouble k, gdouble *xs, gint xstride, gdouble *src,
                       gint sstride, gdouble *normals, gint nstr,
                       gdouble *dipoles, gint dstr, gint nq, gint nsrc,
                       gdouble *xf, gdouble *field, gint fstr);
gint wbfmm_total_field_grad(gdouble k, gdouble *xs, gint xstride,
                            gdouble *src, gint sstride, gdouble *normals,
                            gint nstr, gdouble *dipoles, gint dstr, gint nq,
                            gint nsrc, gdouble *xf, gdouble *field,
                            gint fstr);

/* src/tree.c:86 */
wbfmm_target_list_t *wbfmm_target_list_new(wbfmm_tree_t *t, guint npts);
/* src/targets.c:87: targets are located and sorted on the GPU; the
   downward pass then skips boxes with neither sources nor targets */
gint wbfmm_target_list_add_points(wbfmm_target_list_t *l, gpointer pts,
                                  gsize pstr, guint npts);
/* src/wbfmm.h:821-822 */
gint wbfmm_target_list_coefficients_init(wbfmm_target_list_t *l,
                                         guint field);
/* src/targets.c:168-422 (Laplace precompute).  The GPU evaluates the
   regular basis at apply time, so nothing is stored per target. */
gint wbfmm_laplace_target_list_local_coefficients(wbfmm_target_list_t *l,
                                                  guint source,
                                                  gdouble *work);
/* src/wbfmm.h:1099-1101 (Helmholtz precompute): keeps the wavenumber; the
   GPU evaluates the regular functions on the fly */
gint wbfmm_target_list_local_coefficients(wbfmm_target_list_t *l, gdouble k,
                                          gdouble *work);
/* src/targets.c:424-603: fused L2P + near field for every target.  Laplace:
   1, 3 or 4 values per target (potential, gradient, both with the
   potential first); Helmholtz: 2, 6 or 8 doubles (complex).  Gradients and
   POTENTIAL|GRADIENT are evaluated, where the reference asserts
   (src/targets.c:280, :562). */
gint wbfmm_target_list_local_field(wbfmm_target_list_t *l, gdouble *src,
                                   gint sstr, gdouble *d, gint dstr,
                                   gdouble *f, gint fstr);

/* ------------------------------------------------------------------ */
/* additions (no reference counterpart)                               */
/* ------------------------------------------------------------------ */
/* copy the coefficients of one level (which: 0 mps, 1 mpr) to host
   memory in the reference's packing (8-sibling interleave). */
gint wbfmm_tree_coefficients_download(wbfmm_tree_t *t, guint level,
                                      gint which, gdouble *host);
/* whether add_points/refine fill the host-side ip[] and boxes[l][]
   arrays (default TRUE, as the reference does).  FALSE keeps the tree on
   the GPU only, which is what a solve loop with device-resident inputs
   wants. */
gint wbfmm_tree_host_mirror(wbfmm_tree_t *t, gboolean on);
/* fused L2P + near-field for ntgt points (host or device, byte stride
   tstr), field in {POTENTIAL, GRADIENT, POTENTIAL|GRADIENT}: f gets 1, 3
   or 4 values per target at stride fstr (elements), accumulated. */
gint wbfmm_laplace_field_targets(wbfmm_tree_t *t, gdouble *xt, gsize tstr,
                                 guint ntgt, gdouble *src, gint sstr,
                                 gdouble *d, gint dstr, guint field,
                                 gdouble *f, gint fstr);
/* Helmholtz potential and/or gradient at ntgt points (host or device, byte
   stride tstr) in one launch: f gets 2 (potential), 6 (gradient) or 8
   (both, potential first) doubles per target at stride fstr (elements),
   accumulated */
gint wbfmm_helmholtz_field_targets(wbfmm_tree_t *t, gdouble k, gdouble *xt,
                                   gsize tstr, guint ntgt, gdouble *src,
                                   gint sstr, gdouble *d, gint dstr,
                                   guint field, gdouble *f, gint fstr);
/* wait for all GPU work queued by the library */
gint wbfmm_hip_synchronize(void);
/* release a tree and its device memory (the reference has no destroy) */
gint wbfmm_tree_free(wbfmm_tree_t *t);
gint wbfmm_target_list_free(wbfmm_target_list_t *l);
gint wbfmm_shift_operators_free(wbfmm_shift_operators_t *op);
/* drop the Helmholtz coaxial tables (and their wavenumber) of an operator
   set, keeping its rotation tables, so the set serves another wavenumber
   after new coaxial_SR_init / _SS_init calls */
gint wbfmm_shift_operators_coaxial_reset(wbfmm_shift_operators_t *op);
/* precision of a tree's expansions, 64 (default) or 32, set before the
   coefficients are initialised; inputs and outputs stay FP64 and the
   download returns the coefficients in the reference's scaling.
   Laplace, 32: FP32 coefficients scaled by each level's box width and FP32
   matrix-core translations (orders 2..15, the same order at every level);
   Legendre and power recurrences and the near field stay FP64.
   Helmholtz, 32 (mixed precision): levels with k a_l <= 2 (a_l the box
   circumradius; WBFMM_HEL_FP64_KA sets the threshold) and orders 2..27
   hold FP32 coefficients scaled per degree by s_n = min(1, (k a_l)^n /
   (2n+1)!!) and run FP32 matrix-core translations and an FP32 near field
   with compensated sums; the other levels (large k a_l, near the root)
   stay FP64, and passes between the two run in FP64 on converted copies.
   Operator tables and the Bessel/Legendre recurrences stay FP64.  The
   precision of each level is fixed by the wavenumber of the leaf
   expansions. */
gint wbfmm_tree_set_precision(wbfmm_tree_t *t, gint bits);

/* ---- single precision (src/wbfmm.h:432-440, :690-740, :996-1112) ------
   The Laplace tree pipeline, its target lists and the direct sum with
   float arguments.  A tree made by wbfmm_tree_new_f is an FP32 tree, as
   after wbfmm_tree_set_precision(t, 32): orders 2..15, the same order at
   every level.  Float arrays are host arrays; float outputs are
   accumulated in float.  The public origin (wbfmm_tree_origin) holds
   three gfloat, as in the reference; the passes use a double copy. */
wbfmm_tree_t *wbfmm_tree_new_f(gfloat *x, gfloat D, guint maxpoints);
gint wbfmm_tree_add_points_f(wbfmm_tree_t *t, gpointer pts, gsize pstr,
                             gpointer normals, gsize nstr, guint npts,
                             gboolean sorted);
/* src/wbfmm.h:1011-1013: float records (host arrays), codes from the
   coordinates widened to double */
gint wbfmm_tree_sort_points_f(wbfmm_tree_t *t, gpointer pts, gsize psize, guint npts);
gint wbfmm_tree_refine_f(wbfmm_tree_t *t);
gint wbfmm_tree_laplace_coefficient_init_f(wbfmm_tree_t *t, guint l, guint nr,
                                           guint ns);
gint wbfmm_tree_laplace_leaf_expansions_f(wbfmm_tree_t *t, gfloat *src,
                                          gint sstr, gfloat *dipoles,
                                          gint dstr, gboolean zero_expansions,
                                          gfloat *work);
gint wbfmm_shift_angle_table_init_f(void);
wbfmm_shift_operators_t *wbfmm_shift_operators_new_f(guint L, gboolean bw,
                                                     gfloat *work);
gint wbfmm_laplace_coaxial_translate_init_f(gint N);
gint wbfmm_laplace_upward_pass_f(wbfmm_tree_t *t, wbfmm_shift_operators_t *op,
                                 guint level, gfloat *work);
gint wbfmm_laplace_downward_pass_ref_f(wbfmm_tree_t *t,
                                       wbfmm_shift_operators_t *op,
                                       guint level, gfloat *work,
                                       gint nthreads);
#define wbfmm_laplace_downward_pass_f(_t, _op, _level, _work, _nth) \
  wbfmm_laplace_downward_pass_ref_f(_t, _op, _level, _work, _nth)
gint wbfmm_laplace_box_field_f(wbfmm_tree_t *t, guint level, guint b,
                               gfloat *src, gint sstr, gfloat *d, gint dstr,
                               guint field, gboolean eval_neighbours,
                               gfloat *x, gfloat *f, gint fstr, gfloat *work);
gint wbfmm_laplace_field_direct_f(gfloat *xs, gint xstride, gfloat *n,
                                  gint nstr, gint nsrc, gfloat *src,
                                  gint sstr, gfloat *d, gint dstr, gint nq,
                                  guint field, gfloat *xf, gfloat *f,
                                  gint fstr);
wbfmm_target_list_t *wbfmm_target_list_new_f(wbfmm_tree_t *t, guint npts);
gint wbfmm_target_list_add_points_f(wbfmm_target_list_t *l, gpointer pts,
                                    gsize pstr, guint npts);
gint wbfmm_laplace_target_list_local_coefficients_f(wbfmm_target_list_t *l,
                                                    guint source,
                                                    gfloat *work);
gint wbfmm_target_list_local_field_f(wbfmm_target_list_t *l, gfloat *src,
                                     gint sstr, gfloat *d, gint dstr,
                                     gfloat *f, gint fstr);

/* The Helmholtz pipeline with float arguments (src/wbfmm.h:938-951,
   :997-1003, :1059-1082, :1102-1104).  A tree made by wbfmm_tree_new_f
   runs the mixed-precision Helmholtz path (as after
   wbfmm_tree_set_precision(t, 32)); complex strengths are pairs of floats
   at element stride, float outputs are accumulated in float. */
gint wbfmm_tree_coefficient_init_f(wbfmm_tree_t *t, guint l, guint nr, guint ns);
gint wbfmm_tree_leaf_expansions_f(wbfmm_tree_t *t, gfloat k, gfloat *src, gint sstr,
                                  gfloat *dipoles, gint dstr, gboolean zero_expansions,
                                  gfloat *work);
gint wbfmm_shift_operators_coaxial_SR_init_f(wbfmm_shift_operators_t *w, gfloat D,
                                             guint level, guint L, gfloat k, gfloat *work);
gint wbfmm_shift_operators_coaxial_SS_init_f(wbfmm_shift_operators_t *w, gfloat D,
                                             guint level, guint L, gfloat k, gfloat *work);
gint wbfmm_upward_pass_f(wbfmm_tree_t *t, wbfmm_shift_operators_t *op, guint level,
                         gfloat *work);
gint wbfmm_downward_pass_ref_f(wbfmm_tree_t *t, wbfmm_shift_operators_t *op, guint level,
                               gfloat *work, gint nthreads);
#define wbfmm_downward_pass_f(_t, _op, _level, _work, _nth) \
  wbfmm_downward_pass_ref_f(_t, _op, _level, _work, _nth)
gint wbfmm_tree_box_local_field_f(wbfmm_tree_t *t, guint level, guint b, gfloat k,
                                  gfloat *x, gfloat *f, gint fstr, gfloat *src, gint sstr,
                                  gfloat *d, gint dstr, gboolean eval_neighbours,
                                  guint field, gfloat *work);
gint wbfmm_target_list_local_coefficients_f(wbfmm_target_list_t *l, gfloat k,
                                            gfloat *work);
gint wbfmm_total_field_f(gfloat k, gfloat *xs, gint xstride, gfloat *src, gint sstride,
                         gfloat *normals, gint nstr, gfloat *dipoles, gint dstr, gint nq,
                         gint nsrc, gfloat *xf, gfloat *field, gint fstr);
gint wbfmm_total_field_grad_f(gfloat k, gfloat *xs, gint xstride, gfloat *src,
                              gint sstride, gfloat *normals, gint nstr, gfloat *dipoles,
                              gint dstr, gint nq, gint nsrc, gfloat *xf, gfloat *field,
                              gint fstr);

#ifdef __cplusplus
}
#endif

#endif /* WBFMM_H_INCLUDED */
