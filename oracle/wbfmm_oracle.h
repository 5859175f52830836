/*
 * wbfmm_oracle.h -- CPU restatement of the octree FMM evaluation path of
 * mjcarley/wbfmm, used ONLY as a test oracle (pytest, the smoke check and
 * the CPU-baseline leg of bench.py).  Nothing in the product library links
 * or calls it.
 *
 * Conventions follow the reference:
 *   - Legendre functions with the Gumerov-Duraiswami normalisation
 *     (src/util.c:67-116);
 *   - Laplace coefficients real-packed: (n,0) at n^2, (n,m>0) re/im at
 *     n^2+2m-1 / n^2+2m (src/wbfmm-private.h:76);
 *   - Helmholtz coefficients complex, index n(n+1)+m (src/wbfmm.h:242);
 *   - octree Morton index with 20 bits per axis, x in bit 0, y in bit 1,
 *     z in bit 2 of each triple (src/indexing.c:68-163, src/boxes.c:324-339).
 *
 * All arrays are dense and box-major inside the oracle (its own layout);
 * tests convert to the reference's sibling-interleaved layout where needed.
 */

#ifndef WBFMM_ORACLE_H
#define WBFMM_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- special functions ------------------------------------------------ */
void orc_legendre(int N, double C, double S, double *P);
void orc_bessel_j(int N, double x, double *j);
void orc_bessel_h(int N, double x, double *h);
int  orc_truncation_number(double k, double D, int level, double tol);

/* ---- Morton indexing --------------------------------------------------- */
uint64_t orc_morton_encode(uint32_t x, uint32_t y, uint32_t z);
void     orc_morton_decode(uint64_t m, uint32_t *x, uint32_t *y, uint32_t *z);
uint64_t orc_point_index(const double *x, const double *x0, double D);
int      orc_box_neighbours(int level, uint64_t idx, uint64_t *nb);
int      orc_box_list4(int level, uint64_t idx, uint64_t *list, int *off);

/* ---- rotations --------------------------------------------------------- */
int  orc_rotation_size(int N);
void orc_H_rotation(double *H, int N, double th);
void orc_shift_angles(int dx, int dy, int dz, double *th, double *ph,
                      double *ch, double *r);

/* ---- Laplace operators ------------------------------------------------- */
void orc_lap_p2m(int N, const double *x0, const double *xs, double q,
                 double *M);
void orc_lap_p2m_normal(int N, const double *x0, const double *xs,
                        const double *nrm, double d, double *M);
double orc_lap_m_eval(int N, const double *x0, const double *M,
                      const double *xf);
void orc_lap_l_eval(int N, const double *x0, const double *L,
                    const double *xf, int field, double *f);
/* second derivatives (xx, yy, zz, xy, yz, zx) of a local expansion at xf,
   added into h[0..5] */
void orc_lap_l_hess(int N, const double *x0, const double *L, const double *xf, double *h);
double orc_lap_coaxial_cfft(int kind, int n, int nu, int m);
void orc_lap_coaxial(int kind, double *Co, int No, const double *Ci, int Ni,
                     double t, double sc);
void orc_lap_rotate(double *Co, const double *Ci, int N, const double *H,
                    double ph, double ch, double sc);

/* ---- Helmholtz operators ----------------------------------------------- */
void orc_hel_p2m(double k, int N, const double *x0, const double *xs,
                 const double *q, double *M);
void orc_hel_p2m_normal(double k, int N, const double *x0, const double *xs,
                        const double *nrm, const double *d, double *M);
void orc_hel_h_eval(double k, int N, const double *x0, const double *M,
                    const double *xf, double *f);
void orc_hel_j_eval(double k, int N, const double *x0, const double *L,
                    const double *xf, int field, double *f);
void orc_hel_coaxial_table(int kind, int L, double kr, double *E);
void orc_hel_coaxial(double *Co, int No, const double *Ci, int Ni,
                     const double *E, int L, int is_complex, double sc);
void orc_hel_rotate(double *Co, const double *Ci, int N, const double *H,
                    double ph, double ch, double sc);

/* ---- FP64 direct sums (reference src/direct.c, src/direct-laplace.c) ---- */
void orc_lap_direct(int nsrc, const double *xs, const double *q,
                    const double *nrm, const double *d,
                    int ntgt, const double *xt, int field, double cutoff,
                    double *f);
void orc_hel_direct(double k, int nsrc, const double *xs, const double *q,
                    const double *nrm, const double *d,
                    int ntgt, const double *xt, int field, double cutoff,
                    double *f);

/* ---- full solves, following the reference drivers ---------------------- */
typedef struct {
  int problem;          /* 1 Laplace, 2 Helmholtz */
  double k;             /* wavenumber (Helmholtz) */
  double x0[3], D;      /* tree origin and width */
  int depth;
  int order_s[17], order_r[17];
  int field;            /* 1 potential, 2 gradient, 3 both */
  int nthreads;         /* threads for M2L (<=1: serial) */
  int dump_level;       /* if >0, copy mps/mpr of this level out */
  double *dump_mps, *dump_mpr;  /* dense box-major, caller allocated */
  double t_phase[6];    /* seconds: tree, P2M, M2M, M2L+L2L, L2P+P2P, total */
} orc_fmm_cfg;

int orc_fmm(orc_fmm_cfg *cfg, int nsrc, const double *xs, const double *q,
            const double *nrm, const double *d, int ntgt, const double *xt,
            double *f);

#ifdef __cplusplus
}
#endif

#endif
