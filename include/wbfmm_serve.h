/*
 * wbfmm_serve.h -- serving engine over the drop-in API: independent FMM
 * solves (Laplace or Helmholtz, each with its own points, strengths,
 * targets and accuracy) queued by any number of host threads and run side
 * by side on a pool of HIP streams.
 *
 * No counterpart in the reference, which solves one problem per process
 * (tools/wbfmm-tree-calc.c.in, tools/wbfmm-laplace-tree-calc.c.in).  Each
 * request runs the reference's call sequence for its problem (tree,
 * refine, coefficients, leaf expansions, upward and downward passes, field
 * at the targets) on one stream; the engine chooses the tree domain, the
 * depth and the expansion orders unless the request fixes them.
 *
 * Scheduling: requests are taken in arrival order by the first free
 * stream (continuous batching: a finished solve frees its stream and its
 * device memory for the next request at once).  A request is admitted only
 * while the estimated device memory of all running requests, its own
 * included, stays within the budget.  The estimate counts the coefficient
 * arrays, points, targets and field, the largest transient buffer of a
 * pass and the operator tables of the request's order.  A request whose
 * estimate exceeds the whole budget is refused at submission
 * (WBFMM_HIP_ERR_NOMEM).
 *
 * Failures: a solve that fails on the device (an allocation refused, a
 * launch error) or that the drop-in API would abort on ends with a
 * negative status in its result; its device objects are released and the
 * engine goes on with the other requests.
 *
 * Results are deterministic: a request gives the same bits whether it
 * runs alone or among others.  All calls return 0 (or a ticket >= 0) on
 * success and a negative WBFMM_HIP_ERR_* status otherwise.
 */

#ifndef WBFMM_SERVE_H_INCLUDED
#define WBFMM_SERVE_H_INCLUDED

#include "wbfmm.h"
#include "wbfmm_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct wbfmm_serve_s wbfmm_serve_t;

/* One solve.  Arrays may be host or device memory; they are borrowed and
   must stay valid until wbfmm_serve_wait returns for the request. */
typedef struct {
  gint problem;          /* WBFMM_PROBLEM_LAPLACE or WBFMM_PROBLEM_HELMHOLTZ */
  gint precision;        /* 64: FP64 throughout; 32: FP32 expansions and
                            translations (Laplace, orders 2..15); 0: 32
                            where allowed and tol >= 1e-5, else 64 */
  gdouble k;             /* wavenumber (Helmholtz) */
  gdouble tol;           /* relative accuracy that picks the orders */
  guint nsrc;
  const gdouble *xs;     /* source positions, byte stride pstr */
  gsize pstr;
  const gdouble *q;      /* monopole strengths, contiguous: 1 double each
                            (Laplace) or 2 (Helmholtz, re and im); NULL: none */
  const gdouble *normals;/* dipole normals, byte stride nstr; NULL: none */
  gsize nstr;
  const gdouble *dip;    /* dipole strengths, laid out like q */
  guint ntgt;
  const gdouble *xt;     /* target positions, byte stride tstr; NULL: the
                            sources are the targets */
  gsize tstr;
  guint field;           /* WBFMM_FIELD_POTENTIAL and/or WBFMM_FIELD_GRADIENT */
  gdouble *f;            /* output, accumulated (as in the reference) */
  gint fstr;             /* output stride in doubles: >= 1, 3, 4 (Laplace:
                            potential, gradient, both with the potential
                            first) or 2, 6, 8 (Helmholtz, complex) */
  gint depth;            /* 0: chosen by the engine */
  gint order;            /* 0: from tol (every level the same for Laplace;
                            per level by the truncation number for
                            Helmholtz) */
  guint check;           /* targets checked against the FP64 direct sum
                            (0: none) */
} wbfmm_serve_request_t;

typedef struct {
  gint status;           /* 0, or the negative status of the solve */
  gint depth;            /* depth used */
  gint order;            /* largest expansion order used */
  gint precision;
  gdouble queue_ms;      /* submission to start */
  gdouble tables_ms;     /* operator tables built for this request */
  gdouble solve_ms;      /* start to the field written (tables included) */
  gdouble latency_ms;    /* submission to the field written */
  gdouble check_ms;      /* the direct-sum check, after latency_ms */
  gsize footprint;       /* device bytes the request was admitted with */
  gdouble err_potential; /* relative l2 error of the checked targets */
  gdouble err_gradient;  /* against the FP64 direct sum; -1 unchecked */
} wbfmm_serve_result_t;

typedef struct {
  guint64 completed;
  guint64 targets;       /* targets of the completed requests */
  gdouble p50_ms, p99_ms, mean_ms, max_ms;  /* latency_ms percentiles */
  gdouble span_s;        /* first submission to last completion */
  gsize peak_bytes;      /* largest sum of admitted footprints */
} wbfmm_serve_stats_t;

/* nstreams solves at a time (1..64) within budget bytes of device memory
   (0: 90% of the free memory at creation) */
wbfmm_serve_t *wbfmm_serve_new(gint nstreams, gsize budget);
/* queue a request; returns its ticket */
gint64 wbfmm_serve_submit(wbfmm_serve_t *s, const wbfmm_serve_request_t *r);
/* wait for a ticket's result (each ticket can be waited for once) */
gint wbfmm_serve_wait(wbfmm_serve_t *s, gint64 ticket,
                      wbfmm_serve_result_t *res);
/* latency statistics of the requests completed since the last reset */
gint wbfmm_serve_stats(wbfmm_serve_t *s, wbfmm_serve_stats_t *st);
gint wbfmm_serve_reset_stats(wbfmm_serve_t *s);
/* depth, largest order and estimated device bytes the engine would use */
gint wbfmm_serve_plan(wbfmm_serve_t *s, const wbfmm_serve_request_t *r,
                      gint *depth, gint *order, gsize *footprint);
/* autotuner: solve r (into scratch outputs, as many copies at once as the
   engine has streams, the way a batch shares the GPU) at each of the
   ndepth candidate depths; at the fastest, time the kernel variants one
   kind at a time (wbfmm_hip_set_variant: the sparse-level M2L tiles, and for
   Helmholtz the field kernel), and keep the fastest depth and variants for
   requests of the same class (problem, precision, order, log2 of the point
   count); returns the depth in *best */
gint wbfmm_serve_tune(wbfmm_serve_t *s, const wbfmm_serve_request_t *r,
                      const gint *depths, gint ndepth, gint *best);
/* the autotuner's choice for r's class: depth (0 when not tuned) and the
   variants (-1: the library's choice) */
gint wbfmm_serve_tuned(wbfmm_serve_t *s, const wbfmm_serve_request_t *r, gint *depth,
                       gint *var_m2l, gint *var_field);
/* for failure-handling tests: the next request submitted has its first
   count device allocations refused (as when memory runs out), so that it
   ends with WBFMM_HIP_ERR_NOMEM in its result */
gint wbfmm_serve_debug_fail_allocations(wbfmm_serve_t *s, gint count);
/* finish the queued work and release the engine */
gint wbfmm_serve_free(wbfmm_serve_t *s);

#ifdef __cplusplus
}
#endif

#endif /* WBFMM_SERVE_H_INCLUDED */
