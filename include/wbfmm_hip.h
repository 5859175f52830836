/*
 * wbfmm_hip.h -- thin C ABI between the host layer (C, wbfmm_api.c, which
 * implements the reference API of wbfmm.h) and the hand-written HIP
 * kernels for gfx950.  Plain pointers and sizes only.  Unless stated,
 * pointers are device pointers, every call is asynchronous on the
 * library's stream and returns 0 on success or a negative status.
 *
 * None of these entry points has a one-to-one counterpart in the
 * reference (whose passes are C loops); each comment names the reference
 * routine whose work the kernel performs.
 */

#ifndef WBFMM_HIP_H_INCLUDED
#define WBFMM_HIP_H_INCLUDED

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define WBFMM_HIP_OK 0
#define WBFMM_HIP_ERR_RUNTIME (-1)
#define WBFMM_HIP_ERR_ARG (-2)
#define WBFMM_HIP_ERR_ORDER (-3)
#define WBFMM_HIP_ERR_DOMAIN (-4)
#define WBFMM_HIP_ERR_NOMEM (-5)

/* largest expansion order the kernels accept */
#define WBFMM_HIP_MAX_ORDER 40

/* ---- runtime ---------------------------------------------------------- */
int wbfmm_hip_init(void);
/* device memory from the default pool, ordered on the calling thread's
   stream (no device-wide wait); free returns it when the stream gets
   there */
int wbfmm_hip_malloc(void **p, size_t bytes);
int wbfmm_hip_free(void *p);
/* streams: every call of a thread runs on the library stream unless the
   thread selected one of its own with wbfmm_hip_set_stream (NULL selects
   the library stream again; *prev, when not NULL, gets the old choice).
   Objects (trees, target lists, operators) are used by one stream at a
   time; separate objects may run concurrently on separate streams. */
int wbfmm_hip_stream_create(void **s);
int wbfmm_hip_stream_destroy(void *s);
int wbfmm_hip_set_stream(void *s, void **prev);
void *wbfmm_hip_get_stream(void);
/* kernel variants of the calling thread (no counterpart in the reference:
   the serving engine's autotuner times them per request class and sets
   the winner around each solve).  value -1 restores the library's own
   choice; other values force a variant where the library would use it:
   - WBFMM_VARIANT_M2L, at sparse levels (fewer than 60 % of the boxes
     hold targets): 0 chunks of targets running every offset, 1 tiles of
     one offset (Laplace FP64; Helmholtz: as 2), 2 tiles of one offset
     class (the default);
   - WBFMM_VARIANT_FIELD (Helmholtz L2P + P2P): 0 one thread per target
     (the default), 1 one wavefront per target box with the neighbours'
     sources staged through LDS. */
#define WBFMM_VARIANT_M2L 0
#define WBFMM_VARIANT_FIELD 1
#define WBFMM_VARIANT_COUNT 2
int wbfmm_hip_set_variant(int which, int value);
int wbfmm_hip_get_variant(int which);
/* free and total device memory */
int wbfmm_hip_mem_info(size_t *free_bytes, size_t *total_bytes);
/* copy of any kind (host/device either side), ordered on the stream;
   returns after the copy when a host buffer is involved */
int wbfmm_hip_memcpy(void *dst, const void *src, size_t bytes);
int wbfmm_hip_memset(void *p, int v, size_t bytes);
/* 1 when p is device (or managed) memory, 0 for host memory */
int wbfmm_hip_is_device(const void *p);
int wbfmm_hip_sync(void);
const char *wbfmm_hip_last_error(void);
/* failure-handling tests: the calling thread's next n device allocations
   (wbfmm_hip_malloc and the library's own) return WBFMM_HIP_ERR_NOMEM */
int wbfmm_hip_debug_fail_allocs(int n);

/* per-kernel timing with HIP events on the library stream:
   enable(1) starts recording every launch; query returns, for kernel
   name, the number of launches and the summed milliseconds. */
int wbfmm_hip_timing_enable(int on);
int wbfmm_hip_timing_reset(void);
int wbfmm_hip_timing_query(const char *name, int *launches, double *ms);

/* ---- tree (src/boxes.c:118-277, src/indexing.c) ----------------------- */
/* Morton codes of n points read at byte stride pstr (host or device);
   writes xyz (4 doubles per point), codes and identity indices; *nbad
   (host) gets the number of points outside the open domain cube. */
int wbfmm_hip_morton(const void *pts, size_t pstr, unsigned n,
                     const double *x0, double D, double *xyz,
                     uint64_t *codes, uint32_t *idx, unsigned *nbad);
/* n rows of 3 doubles read at byte stride pstr (host or device) copied
   to 4-double device rows (4th entry 0) */
int wbfmm_hip_stage4(const void *pts, size_t pstr, unsigned n, double *out);
/* stable sort of (codes, idx) pairs by code */
int wbfmm_hip_sort(uint64_t *codes, uint32_t *idx, unsigned n);
/* out[4i..4i+3] = in[4 idx[i] ..] */
int wbfmm_hip_gather4(const double *in, const uint32_t *idx, unsigned n,
                      double *out);
/* record i of out = record idx[i] of in, records of rsize bytes (device
   arrays) */
int wbfmm_hip_gather_records(const void *in, size_t rsize, const uint32_t *idx,
                             unsigned n, void *out);
/* out[i] = src[idx[i]*sstr] (src host or device) */
int wbfmm_hip_gather_strided(const double *src, int sstr, int ncomp,
                             const uint32_t *idx, unsigned n, double *out);
/* dense [start,end) of every box at `level` from sorted codes (both
   arrays 8^level, zeroed here first) */
int wbfmm_hip_box_ranges(const uint64_t *codes, unsigned n, unsigned level,
                         uint32_t *start, uint32_t *end);
/* compact list of the boxes with end>start (or flag!=0 when flag is not
   NULL) in Morton order; *count (host) gets the list length */
int wbfmm_hip_compact_boxes(const uint32_t *start, const uint32_t *end,
                            const uint32_t *flag, unsigned level,
                            uint32_t *list, unsigned *count);
/* coordinate bounds of n points (host or device, byte stride pstr) into
   host lo[3], hi[3]; returns after the reduction */
int wbfmm_hip_bounds(const void *pts, size_t pstr, unsigned n, double *lo,
                     double *hi);
/* dst[i*dstr + c] += src[i*sstr + c], c < ncomp, i < n (device arrays) */
int wbfmm_hip_accumulate(double *dst, int dstr, const double *src, int sstr,
                         int ncomp, unsigned n);
/* flag[b] |= 1 for every box at `level` holding one of the n codes */
int wbfmm_hip_mark_boxes(const uint64_t *codes, unsigned n, unsigned level,
                         uint32_t *flag);

/* ---- operator tables --------------------------------------------------- */
/* Real rotation matrices of every list-4 offset (343 slots), of the 8
   child->parent and 8 parent->child directions, forward and back, up to
   order L, built on the host in FP64 from the Gumerov-Duraiswami
   recursions and uploaded.  Returned as one device block; *layout gets
   the element offsets (see wbfmm_tables.c). */
typedef struct {
  int L;                 /* order of the tables */
  size_t nrot;           /* doubles per rotation: sum_{n<=L} (2n+1)^2 */
  size_t m2l_fwd, m2l_back;   /* 343 * nrot each */
  size_t up_fwd, up_back;     /* 8 * nrot each (child -> parent) */
  size_t dn_fwd, dn_back;     /* 8 * nrot each (parent -> child) */
  size_t total;
  double r[343];         /* offset length in box widths (0: not list 4) */
} wbfmm_rot_layout_t;
int wbfmm_hip_rotation_tables(int L, double **dev, wbfmm_rot_layout_t *lay);
/* Laplace coaxial coefficient tables (unit distance), kinds SS, RR, SR,
   c[kind][(m*(L+1)+n)*(L+1)+nu], closed forms of src/laplace.c:542-725 */
int wbfmm_hip_laplace_coaxial_tables(int L, double **dev);

/* Plan of the MFMA form of the Laplace M2L at order N.
   Each translation (rotate, coaxial S|R, rotate back) is split as
     rotate  = Zout(phi) H(theta) Zin(chi),  phi = -pi/2 fixed,
   so the matrices depend on the polar angle theta only (49 values over
   the list-4 offsets) and the azimuthal factors Zin/Zout (2x2 rotations
   of each (re, im) pair by m chi) are applied on the vector units.  With
   phi fixed, each real degree block of H splits into a cos part
   ((n+1)x(n+1)) and a sin part (n x n); the coaxial S|R shift is one
   (N+1-m)^2 block per order m shared by both parts.  The blocks of each
   stage are bin-packed (first fit, decreasing) into NB bins of at most 16
   rows, each bin one 16 x (4 nk) MFMA tile (v_mfma_f64_16x16x4f64).
   Slots: rows and k positions of bin b are slots 16 b + i.
     S = max(16 NB, 2 NU) slots carry data between stages, slot S is a
       dummy destination for padding rows;
     map1/map2/map3[slot]: destination slot of a stage's output row in the
       next stage's k order (map3: tr(n,nu) -> u, ti(n,nu) -> NU + u with
       u = n(n+1)/2 + nu), shifted left by 6, plus the degree n (map1) or
       n+1 (map2) of the t^-k factor the coaxial tables carry (padding
       rows: slot S);
     s1[u]: stage-1 k slots of the rotated re and im parts of unit
       u = (n, m) (low and high 16 bits, high = S for m = 0);
     th/rix/chi/r: theta-table index, distance index, azimuth and length
       (box widths) of each list-4 offset o = (dx+3)*49 + (dy+3)*7 + dz+3
       (dx = source - target); rr[NR]: the distinct lengths (15).
   Tables (MFMA lane order: lane l holds A[l % 16][4 j + l / 16]):
     A1, A3: NTH x KT x 64 (stages 1 and 3, per theta);
     A2: NR x KT x 64, the coaxial S|R at distance rr[i] w for one box
       width w (wbfmm_hip_lap_m2l_coax);
     CS: 343 x (N+1) x 2, cos and sin of m chi for each offset. */
#define WBFMM_M2L_MAXN 30
#define WBFMM_M2L_MAXB 96
#define WBFMM_M2L_MAXU 528
/* bins: 16-row groups of the blocks of a stage.  Blocks of up to 16 rows
   share bins (first fit, decreasing); a larger block spans consecutive
   bins of its own.  Bin b's k-steps j read input rows kbase[b] + 4j ..
   +3: kbase = 16b for shared bins, the block's first row for the bins of
   a large block, whose every row tile reads all the block's columns. */
typedef struct {
  int N, C, NU, NB, KT, NTH, S, NR;
  int nk[WBFMM_M2L_MAXB], koff[WBFMM_M2L_MAXB], kbase[WBFMM_M2L_MAXB];
  int map1[16 * WBFMM_M2L_MAXB], map2[16 * WBFMM_M2L_MAXB],
      map3[16 * WBFMM_M2L_MAXB];
  int s1[WBFMM_M2L_MAXU];
  int th[343], rix[343];
  double chi[343];
  double r[343];
  double rr[32];
} wbfmm_m2l_plan_t;
/* builds the plan of order N (1 <= N <= 24) and, for the pointers that
   are not NULL, fills the host
   tables A1, A3 (NTH*KT*64 doubles each) and CS (343*(N+1)*2); call once
   with NULL tables to learn the sizes */
int wbfmm_hip_lap_m2l_tables(int N, wbfmm_m2l_plan_t *plan, double *A1,
                             double *A3, double *CS);
/* host coaxial table A2 (NR*KT*64 doubles) of a plan for box width w */
int wbfmm_hip_lap_m2l_coax(const wbfmm_m2l_plan_t *plan, double w,
                           double *A2);
/* plan and tables of order N uploaded to the device (plan_dev: device copy
   of *plan) */
int wbfmm_hip_lap_m2l_plan(int N, wbfmm_m2l_plan_t *plan, void **plan_dev,
                           double **A1, double **A3, double **CS);
/* MFMA M2L for target boxes grouped by Morton child digit: cls holds the
   targets of class c at [cstart[c], cstart[c]+ccount[c]); 16 targets of
   one class per wavefront, accumulating into mpr.  Expansions are stored
   box-major (box b at b*C).  sstart/send: dense point ranges of the
   level's boxes (empty boxes are skipped); A2: device coaxial table for
   this level's box width D/2^level. */
int wbfmm_hip_lap_m2l_mfma(const uint32_t *cls, const unsigned *cstart,
                           const unsigned *ccount, const uint32_t *sstart,
                           const uint32_t *send, unsigned level, double D,
                           int N, const double *mps, double *mpr,
                           const void *plan_dev, const double *A1,
                           const double *A3, const double *A2,
                           const double *CS,
                           const wbfmm_m2l_plan_t *plan_host);
/* wbfmm_hip_lap_m2l_mfma without any device allocation: no counterpart
   in the reference; for callers that cannot allocate device memory (code
   under HIP graph capture, tight serving loops).  Same arguments, kernels
   and results as wbfmm_hip_lap_m2l_mfma; its scratch (zero expansion,
   occupancy bitmap, split partials) is a 64 MiB device array of the
   library, one per GPU, shared by every call on that GPU, so calls that
   overlap on one GPU race.  Refused with WBFMM_HIP_ERR_ARG and a message
   (wbfmm_hip_last_error) before any launch: levels above 8, and levels
   under 1024 chunks of 16 targets whose split partials
   (blocks per chunk x targets x (N+1)^2 x 8 bytes) exceed 62 MiB. */
int wbfmm_hip_lap_m2l_mfma_noalloc(const uint32_t *cls, const unsigned *cstart,
                                   const unsigned *ccount, const uint32_t *sstart,
                                   const uint32_t *send, unsigned level, double D,
                                   int N, const double *mps, double *mpr,
                                   const void *plan_dev, const double *A1,
                                   const double *A3, const double *A2,
                                   const double *CS,
                                   const wbfmm_m2l_plan_t *plan_host);
/* tables of the MFMA M2M / L2L for plan P (same bins and slot maps):
   A1, A3: 2 x KT x 64 (polar angles acos(1/sqrt3), acos(-1/sqrt3));
   CSup, CSdn: 8 x (N+1) x 2, cos and sin of m chi of child c's
   child -> parent and parent -> child directions */
int wbfmm_hip_lap_shift_tables(const wbfmm_m2l_plan_t *plan, double *A1,
                               double *A3, double *CSup, double *CSdn);
/* stage-2 tiles (KT x 64): kind 0 S|S c (-t)^(n-nu) (M2M), kind 1 R|R
   c t^(nu-n) (L2L), t = sqrt(3)/2 times the child box width */
int wbfmm_hip_lap_shift_coax(const wbfmm_m2l_plan_t *plan, int kind, double t,
                             double *A2);
/* MFMA M2M (mode 0: each listed parent from its children, child ranges
   cstart/cend; the parent is overwritten) or L2L (mode 1: each listed
   parent's local expansion added to its children with cflag != 0, all
   when NULL); in/out: source/destination level arrays (box-major), orders
   equal (2..15) (src/shift-laplace.c:225-789) */
int wbfmm_hip_lap_shift_mfma(int mode, const uint32_t *parents, unsigned npar,
                             const uint32_t *cstart, const uint32_t *cend,
                             const uint32_t *cflag, int N, const double *in,
                             double *out, const void *plan_dev,
                             const double *A1, const double *A3,
                             const double *A2, const double *CS,
                             const wbfmm_m2l_plan_t *plan_host);
/* one v_mfma_f64_16x16x4f64 with A[i][k] = 1+i+100k, B[k][j] = 1+3k+1000j;
   host_out (256 doubles) gets the four result values of each lane */
int wbfmm_hip_mfma_layout_probe(double *host_out);
/* the same product with v_mfma_f32_16x16x4f32 (256 floats) */
int wbfmm_hip_mfma_layout_probe_f32(float *host_out);

/* ---- FP32 Laplace passes ------------------------------------------------
   Same work as the FP64 entry points above, with FP32 coefficients and
   FP32 matrix-core products.  The coefficients of each level are scaled
   by its box width w (M_n / w^n, L_n w^(n+1)) so that FP32 holds them at
   every depth; the tables are the FP64 ones in that scaling, rounded to
   FP32: M2L A2 from wbfmm_hip_lap_m2l_coax at w = 1, M2M / L2L A2 from
   wbfmm_hip_lap_shift_coax_unit (the same at every level).  Legendre and
   power recurrences (P2M, L2P) run in FP64, and so does the near field. */
int wbfmm_hip_lap_shift_coax_unit(const wbfmm_m2l_plan_t *plan, int kind, double *A2);
int wbfmm_hip_lap_m2l_mfma_f32(const uint32_t *cls, const unsigned *cstart,
                               const unsigned *ccount, const uint32_t *sstart,
                               const uint32_t *send, unsigned level, int N,
                               const float *mps, float *mpr, const void *plan_dev,
                               const float *A1, const float *A3, const float *A2,
                               const float *CS, const wbfmm_m2l_plan_t *plan_host);
int wbfmm_hip_lap_shift_mfma_f32(int mode, const uint32_t *parents, unsigned npar,
                                 const uint32_t *cstart, const uint32_t *cend,
                                 const uint32_t *cflag, int N, const float *in,
                                 float *out, const void *plan_dev, const float *A1,
                                 const float *A3, const float *A2, const float *CS,
                                 const wbfmm_m2l_plan_t *plan_host);
int wbfmm_hip_lap_p2m_f32(const uint32_t *boxes, unsigned nbox,
                          const uint32_t *start, const uint32_t *end,
                          const double *xyz, const double *q, const double *nrm,
                          const double *dip, unsigned level, const double *x0,
                          double D, int N, float *mps);
int wbfmm_hip_lap_field_f32(const double *xt, const uint32_t *tid, unsigned n,
                            const uint32_t *start, const uint32_t *end,
                            const double *xyz, const double *q, const double *nrm,
                            const double *dip, unsigned level, const double *x0,
                            double D, int N, const float *mpr, int field,
                            double *f, int fstr);
/* Helmholtz M2L on matrix cores (Laplace plan of the same order, in the
   basis P_m = x'_m + x'_{-m}, Q_m = x'_m - x'_{-m} of the phase-rotated
   expansion): rotation tiles of the plain H(theta) (NTH x KT x 64, both
   rotations) and coaxial tiles from nr host S|R tables E of order L >= N
   (nr x 2 x KT x 64: real, imaginary part) */
int wbfmm_hip_hel_m2l_rot_tiles(const wbfmm_m2l_plan_t *plan, double *A);
/* the same tiles at nth given polar angles (nth x KT x 64); the M2M / L2L
   use the child directions' acos(1/sqrt 3) (index 0) and acos(-1/sqrt 3) */
int wbfmm_hip_hel_rot_tiles(const wbfmm_m2l_plan_t *plan, int nth, const double *theta,
                            double *A);
int wbfmm_hip_hel_m2l_coax_tiles(const wbfmm_m2l_plan_t *plan, int L, int nr,
                                 const double *E, double *A2);
/* the same tiles with each output degree n scaled by sout[n] and each
   input degree l by sin[l] (NULL: 1): the operators of coefficients held
   as M_n / s_n (multipoles) or L_n s_n (local expansions) */
int wbfmm_hip_hel_coax_tiles_scaled(const wbfmm_m2l_plan_t *plan, int L, int nr,
                                    const double *E, const double *sout, const double *sin,
                                    double *A2);
/* out[i] = (float)in[i] on the device */
int wbfmm_hip_to_f32(const double *in, float *out, size_t n);
/* sort a box list by Morton child digit (stable); returns the per-class
   starts and counts on the host */
int wbfmm_hip_split_classes(const uint32_t *list, unsigned n, uint32_t *out,
                            unsigned *cstart, unsigned *ccount);

/* ---- Laplace passes ------------------------------------------------------ */
/* P2M into mps at `level` for the listed leaf boxes (src/laplace.c:854-989) */
int wbfmm_hip_lap_p2m(const uint32_t *boxes, unsigned nbox,
                      const uint32_t *start, const uint32_t *end,
                      const double *xyz, const double *q,
                      const double *nrm, const double *dip,
                      unsigned level, const double *x0, double D, int N,
                      double *mps);
/* the same expansions with one wavefront per box, for boxes holding many
   points (the API picks it at 4 or more points per occupied leaf); _f32:
   into the FP32 level (M_n / w^n) of wbfmm_hip_lap_p2m_f32 */
int wbfmm_hip_lap_p2m_wave(const uint32_t *boxes, unsigned nbox,
                           const uint32_t *start, const uint32_t *end,
                           const double *xyz, const double *q,
                           const double *nrm, const double *dip,
                           unsigned level, const double *x0, double D, int N,
                           double *mps);
int wbfmm_hip_lap_p2m_wave_f32(const uint32_t *boxes, unsigned nbox,
                               const uint32_t *start, const uint32_t *end,
                               const double *xyz, const double *q,
                               const double *nrm, const double *dip,
                               unsigned level, const double *x0, double D, int N,
                               float *mps);
/* M2M: parents (list at level-1) from children at level
   (src/pass-laplace.c:433-486, src/shift-laplace.c:225-442) */
int wbfmm_hip_lap_m2m(const uint32_t *parents, unsigned npar,
                      const uint32_t *cstart, const uint32_t *cend,
                      unsigned level, double D, int Nc, int Np,
                      const double *mps_child, double *mps_parent,
                      const double *rot, const wbfmm_rot_layout_t *lay,
                      const double *coax);
/* M2L for the listed target boxes at `level`, accumulating into mpr
   (src/pass-laplace.c:35-408) */
int wbfmm_hip_lap_m2l(const uint32_t *targets, unsigned ntgt,
                      const uint32_t *sstart, const uint32_t *send,
                      unsigned level, double D, int Ns, int Nr,
                      const double *mps, double *mpr, const double *rot,
                      const wbfmm_rot_layout_t *lay, const double *coax);
/* L2L from the listed parents at `level` into their children with
   cflag != 0 (all children when cflag is NULL) (src/shift-laplace.c:
   562-789) */
int wbfmm_hip_lap_l2l(const uint32_t *parents, unsigned npar,
                      const uint32_t *cflag, unsigned level, double D,
                      int Np, int Nc, const double *mpr_parent,
                      double *mpr_child, const double *rot,
                      const wbfmm_rot_layout_t *lay, const double *coax);
/* L2P + near field for n targets (xyz: 4 doubles each, device) writing
   field values to f[tid[i]*fstr + ...] (accumulated); field bits 1
   potential, 2 gradient; layout: potential first when both are asked
   (src/local-laplace.c:333-630, :1091-1336) */
int wbfmm_hip_lap_field(const double *xt, const uint32_t *tid, unsigned n,
                        const uint32_t *start, const uint32_t *end,
                        const double *xyz, const double *q,
                        const double *nrm, const double *dip,
                        unsigned level, const double *x0, double D, int N,
                        const double *mpr, int field, double *f, int fstr);
/* vector sources, nq components (src/local-laplace.c:333-1336, the curl
   and curl-gradient of :632-1291): L2P + near field for n staged targets
   (rows tid[i] of f, fstr doubles apart).  Component q's local expansions
   are plane q of mpr (plane_stride doubles apart, box-major), its sorted
   strengths q[s nq + q] (monopoles) and dip[s nq + q] (normal dipoles,
   normals nrm).  mode: bit 1 the nq potentials, bit 2 their gradients
   (d_j A_q at 3q + j after the potentials); or bit 4 the curl of the
   first three components (3 values) and, with bit 2, its gradient
   (d_j u_i at 3 + 3j + i; 12 values).  nq <= 8; the curl modes need
   nq >= 3 and take monopoles only, as the reference's. */
int wbfmm_hip_lap_field_vec(const double *xt, const uint32_t *tid, unsigned n,
                            const uint32_t *start, const uint32_t *end, const double *xyz,
                            const double *q, const double *nrm, const double *dip, int nq,
                            unsigned level, const double *x0, double D, int N,
                            const double *mpr, size_t plane_stride, int mode, double *f,
                            int fstr);
/* FP64 direct sums with vector sources (src/direct-laplace.c:87-336):
   the modes and rows of wbfmm_hip_lap_field_vec for every target t (row t
   of f) over every source (strengths nq per source in source order),
   skipping |x - y| <= cutoff */
int wbfmm_hip_lap_direct_vec(const double *xyz, const double *q, const double *nrm,
                             const double *dip, int nq, unsigned nsrc, const double *xt,
                             unsigned ntgt, int mode, double cutoff, double *f, int fstr);
/* FP64 direct sum on the GPU (src/direct-laplace.c:200-336): sources in
   xyz (4 doubles each), q and/or (nrm,dip); cutoff: pairs with r<=cutoff
   skipped (0: none); f[i*fstr + ...] accumulated as in _lap_field */
int wbfmm_hip_lap_direct(const double *xyz, const double *q,
                         const double *nrm, const double *dip,
                         unsigned nsrc, const double *xt, unsigned ntgt,
                         int field, double cutoff, double *f, int fstr);

/* ---- Helmholtz tables (host, wbfmm_tables.c) --------------------------- */
/* geometry of the 343 list-4 offset slots (dx = source - target, slot
   (dx+3)*49 + (dy+3)*7 + dz+3): index of the polar angle in thetas[] (49
   distinct), azimuth chi (0 on the z axis, phi = -pi/2 throughout) and
   index of the length in rr[] (15 distinct); thetas needs 64 slots, rr 16 */
int wbfmm_hip_offset_geometry(double *thetas, int *nth, int *th_idx,
                              double *chi, double *rr, int *nr, int *rix);
/* raw Gumerov-Duraiswami H_n^{nu,m}(theta) (nu = -n..n, m = 0..n, n <= L)
   for nth angles in the reference's index order (src/wbfmm-private.h:
   73-74), uploaded as one device block; *stride gets the doubles per
   angle */
int wbfmm_hip_H_tables(int L, int nth, const double *theta, double **dev,
                       size_t *stride);
/* Helmholtz coaxial coefficients (E|F)_{l n}^m, l <= L, at signed distance
   kr into host E: 2 doubles (re, im) per entry l(l+1)(l+2)/6 + n(n+1)/2 +
   m; kind 0 R|R (= S|S), kind 1 S|R (src/translations.c:205-409) */
int wbfmm_hip_hel_coaxial_table(int kind, int L, double kr, double *E);
/* n tables at kr[0..n-1] on host threads, table i at E + i * 2 ne where
   ne = (L+1)(L+2)(L+3)/6 */
int wbfmm_hip_hel_coaxial_tables(int kind, int L, int n, const double *kr, double *E);

/* ---- Helmholtz passes (FP64, complex coefficients, box-major) ---------- */
/* P2M into mps at `level` for the listed leaf boxes; q, dip: complex
   strengths in sorted order (2 doubles each); nrm: sorted normals, 4
   doubles each (src/boxes.c:341-490) */
int wbfmm_hip_hel_p2m(const uint32_t *boxes, unsigned nbox,
                      const uint32_t *start, const uint32_t *end,
                      const double *xyz, const double *q, const double *nrm,
                      const double *dip, unsigned level, const double *x0,
                      double D, int N, double k, double *mps);
/* M2M: each listed parent from its children (child level ranges
   cstart/cend); Hud: H tables of acos(1/sqrt3) and acos(-1/sqrt3), nrot
   doubles each; SS: R|R table of the child level (src/shift.c:244-401) */
int wbfmm_hip_hel_m2m(const uint32_t *parents, unsigned npar,
                      const uint32_t *cstart, const uint32_t *cend, int Nc,
                      int Np, const double *mps_child, double *mps_parent,
                      const double *Hud, size_t nrot, const double *SS);
/* L2L: each listed parent's local expansion added to its children with
   cflag != 0 (all when NULL); SS: R|R table of the child level
   (src/shift.c:403-585) */
int wbfmm_hip_hel_l2l(const uint32_t *parents, unsigned npar,
                      const uint32_t *cflag, int Np, int Nc,
                      const double *mpr_parent, double *mpr_child,
                      const double *Hud, size_t nrot, const double *SS);
/* M2L on matrix cores for Ns = Nr = N in 2..24 (src/pass.c:56-105,
   :448-537): the targets of level `level' split by Morton child digit
   (cls, cstart, ccount from wbfmm_hip_split_classes), sources by the
   level's box ranges; the Laplace plan of order N (plan_dev on the device,
   plan_host on the host, CS its cos/sin of m chi), the rotation tiles A
   (wbfmm_hip_hel_m2l_rot_tiles) and the level's coaxial tiles A2
   (wbfmm_hip_hel_m2l_coax_tiles), all on the device.  Accumulates into
   mpr; the same sums as wbfmm_hip_hel_m2l to rounding. */
int wbfmm_hip_hel_m2l_mfma(const uint32_t *cls, const unsigned *cstart,
                           const unsigned *ccount, const uint32_t *sstart,
                           const uint32_t *send, unsigned level, int N,
                           const double *mps, double *mpr, const void *plan_dev,
                           const double *A, const double *A2, const double *CS,
                           const wbfmm_m2l_plan_t *plan_host);
/* M2M (mode 0: the occupied children of each listed parent, by
   cstart/cend of the child level, into the parent, overwriting it;
   src/pass.c:581-632) and L2L (mode 1: each listed parent into its
   children flagged in cflag, or all if cflag is NULL, accumulating;
   src/pass.c:550-576) on matrix cores for Nc = Np = N in 2..24: the
   Laplace plan of order N, the rotation tiles A of the child polar angles
   (wbfmm_hip_hel_rot_tiles), the real part A2 of the coaxial tiles of the
   R|R table (wbfmm_hip_hel_m2l_coax_tiles with one table) and the
   child-direction phases CS (CSup for M2M, CSdn for L2L of
   wbfmm_hip_lap_shift_tables), all on the device */
int wbfmm_hip_hel_shift_mfma(int mode, const uint32_t *parents, unsigned npar,
                             const uint32_t *cstart, const uint32_t *cend,
                             const uint32_t *cflag, int N, const double *in, double *out,
                             const void *plan_dev, const double *A, const double *A2,
                             const double *CS, const wbfmm_m2l_plan_t *plan_host);
/* M2L for the listed target boxes, accumulating into mpr; Hth: H tables
   of the offset polar angles; th_idx/chi/rix: per-slot geometry (device);
   SR: one S|R table of ne complex entries per offset length
   (src/pass.c:56-105, :448-537) */
int wbfmm_hip_hel_m2l(const uint32_t *targets, unsigned ntgt,
                      const uint32_t *sstart, const uint32_t *send,
                      unsigned level, int Ns, int Nr, const double *mps,
                      double *mpr, const double *Hth, size_t nrot,
                      const int *th_idx, const double *chi, const int *rix,
                      const double *SR, size_t ne);
/* L2P + near field for n targets (4 doubles each), accumulated into
   f[tid[i]*fstr + ...]: field bit 1 potential (2 doubles), bit 2
   gradient (6), potential first (src/local.c:32-310) */
int wbfmm_hip_hel_field(const double *xt, const uint32_t *tid, unsigned n,
                        const uint32_t *start, const uint32_t *end,
                        const double *xyz, const double *q, const double *nrm,
                        const double *dip, unsigned level, const double *x0,
                        double D, int N, double k, const double *mpr, int field,
                        double *f, int fstr);
/* ---- mixed-precision Helmholtz (FP32 levels) -----------------------------
   An FP32 level holds M_n / s_n (multipoles) and L_n s_n (local
   expansions), s_n the level's per-degree scale (device array of N+2
   doubles), as complex float pairs, box-major.  P2M scales and rounds on
   store; the matrix-core passes are those above on FP32 tiles (operators
   scaled per degree, wbfmm_hip_hel_coax_tiles_scaled); the field kernel
   divides the scale out in FP64. */
int wbfmm_hip_hel_p2m_f32(const uint32_t *boxes, unsigned nbox, const uint32_t *start,
                          const uint32_t *end, const double *xyz, const double *q,
                          const double *nrm, const double *dip, unsigned level,
                          const double *x0, double D, int N, double k, const double *scale,
                          float *mps);
int wbfmm_hip_hel_m2l_mfma_f32(const uint32_t *cls, const unsigned *cstart,
                               const unsigned *ccount, const uint32_t *sstart,
                               const uint32_t *send, unsigned level, int N, const float *mps,
                               float *mpr, const void *plan_dev, const float *A, const float *A2,
                               const float *CS, const wbfmm_m2l_plan_t *plan_host);
int wbfmm_hip_hel_shift_mfma_f32(int mode, const uint32_t *parents, unsigned npar,
                                 const uint32_t *cstart, const uint32_t *cend,
                                 const uint32_t *cflag, int N, const float *in, float *out,
                                 const void *plan_dev, const float *A, const float *A2,
                                 const float *CS, const wbfmm_m2l_plan_t *plan_host);
/* mixed-precision L2P + near field with one thread per target (the form
   of wbfmm_hip_hel_field): FP32 local expansions holding L_n^m s_n
   (scale: the level's s_n, FP64), FP64 regular functions and L2P sums,
   FP32 near-field kernel with compensated sums (src/local.c:32-310) */
int wbfmm_hip_hel_field_f32(const double *xt, const uint32_t *tid, unsigned n,
                            const uint32_t *start, const uint32_t *end, const double *xyz,
                            const double *q, const double *nrm, const double *dip,
                            unsigned level, const double *x0, double D, int N, double k,
                            const float *mpr, const double *scale, int field, double *f,
                            int fstr);
/* L2P + near field with one wavefront per target box (src/local.c:32-310):
   tboxes lists the nbox leaf boxes holding targets, whose sorted targets
   are [tstart[b], tend[b]) of xt (4 doubles each, caller index tid); the
   27 neighbour boxes' sources pass through LDS.  prec 64: FP64 mpr and
   near field; prec 32: FP32 mpr scaled by scale[n], FP32 near field with
   compensated sums.  Output as wbfmm_hip_hel_field. */
int wbfmm_hip_hel_field_box(const uint32_t *tboxes, unsigned nbox, const uint32_t *tstart,
                            const uint32_t *tend, const double *xt, const uint32_t *tid,
                            const uint32_t *start, const uint32_t *end, const double *xyz,
                            const double *q, const double *nrm, const double *dip,
                            unsigned level, const double *x0, double D, int N, double k,
                            int prec, const void *mpr, const double *scale, int field,
                            double *f, int fstr);
/* level conversion for the listed boxes: to32 1: FP64 in -> FP32 out
   (overwritten, or added to when add is 1), to32 0: FP32 in -> FP64 out;
   degree n multiplied by scale[n] */
int wbfmm_hip_hel_convert(int to32, int add, const uint32_t *boxes, unsigned nbox, int N,
                          const void *in, const double *scale, void *out);
/* FP64 Helmholtz direct sum (src/direct.c:35-294); pairs with r <= cutoff
   skipped; same output layout as wbfmm_hip_hel_field */
int wbfmm_hip_hel_direct(const double *xyz, const double *q, const double *nrm,
                         const double *dip, unsigned nsrc, const double *xt,
                         unsigned ntgt, double k, int field, double cutoff,
                         double *f, int fstr);

#ifdef __cplusplus
}
#endif

#endif
