/*
 * points_io.h -- text point files of the reference's driver programs and
 * the drivers' timing stamps (shared by wbfmm-tree-calc.c and
 * wbfmm-laplace-tree-calc.c).
 *
 * File format (tools/wbfmm-laplace-tree-calc.c.in:48-182,
 * tools/wbfmm-tree-calc.c.in:47-136): a header
 *   [number of points] [number of source components] [type code]
 * followed by one line per point:
 *   M   x y z q_0 .. q_{nq-1}                        (monopoles)
 *   MN  x y z q_0 .. q_{nq-1} nx ny nz d_0 .. d_{nq-1} (monopoles, normals,
 *       dipole strengths)
 *   N   x y z nx ny nz d_0 .. d_{nq-1}               (normals and dipole
 *       strengths only; Laplace driver)
 *   F   x y z                                        (field points)
 * Results are written as "x y z f_0 f_1 ..." with %1.16e.
 */

#ifndef WBFMM_POINTS_IO_H
#define WBFMM_POINTS_IO_H

#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

typedef struct {
  int n, nq;         /* points, source components as written in the header */
  int str;           /* doubles per row of xs */
  double *xs;        /* rows of str doubles, coordinates first */
  double *q;         /* monopole strengths (row stride qstr) or NULL */
  double *nrm;       /* normals (row stride nstr) or NULL */
  double *dip;       /* dipole strengths (row stride dstr) or NULL */
  int qstr, nstr, dstr;
} point_file_t;

static int read_point_file(const char *prog, const char *file, int allow_n, point_file_t *p)
{
  FILE *in = file ? fopen(file, "r") : stdin;
  char code[8] = "";
  int i, j;
  memset(p, 0, sizeof(*p));
  if (!in) {
    fprintf(stderr, "%s: cannot open file %s\n", prog, file);
    exit(1);
  }
  if (fscanf(in, "%d %d %7s", &p->n, &p->nq, code) != 3 || p->n < 0 || p->nq < 0) {
    fprintf(stderr, "%s: bad header in %s\n", prog, file ? file : "stdin");
    exit(1);
  }
  fprintf(stderr, "%s: %d point%c of %d components\n", prog, p->n, p->n > 1 ? 's' : ' ', p->nq);
  if (!strcmp(code, "M")) {
    p->str = 3 + p->nq;
  } else if (!strcmp(code, "MN")) {
    p->str = 6 + 2 * p->nq;
  } else if (!strcmp(code, "N") && allow_n) {
    p->str = 6 + p->nq;
  } else if (!strcmp(code, "F")) {
    p->str = 3;
  } else {
    fprintf(stderr, "%s: unknown point type code `%s'\n", prog, code);
    exit(1);
  }
  p->xs = calloc((size_t)p->str * (p->n ? p->n : 1), sizeof(double));
  if (!p->xs) {
    fprintf(stderr, "%s: out of memory\n", prog);
    exit(1);
  }
  for (i = 0; i < p->n; i++)
    for (j = 0; j < p->str; j++)
      if (fscanf(in, "%lg", &p->xs[(size_t)p->str * i + j]) != 1) {
        fprintf(stderr, "%s: short point data in %s\n", prog, file ? file : "stdin");
        exit(1);
      }
  if (file) fclose(in);
  if (!strcmp(code, "M")) {
    p->q = p->xs + 3;
    p->qstr = p->str;
  } else if (!strcmp(code, "MN")) {
    p->q = p->xs + 3;
    p->qstr = p->str;
    p->nrm = p->xs + 3 + p->nq;
    p->nstr = p->str;
    p->dip = p->xs + 6 + p->nq;
    p->dstr = p->str;
  } else if (!strcmp(code, "N")) {
    p->nrm = p->xs + 3;
    p->nstr = p->str;
    p->dip = p->xs + 6;
    p->dstr = p->str;
  }
  return 0;
}

static int parse_origin(double *x, const char *s)
{
  return sscanf(s, "%lg,%lg,%lg", &x[0], &x[1], &x[2]) == 3 ? 0 : -1;
}

static double elapsed(void)
{
  static struct timespec t0;
  static int started = 0;
  struct timespec t;
  if (!started) {
    clock_gettime(CLOCK_MONOTONIC, &t0);
    started = 1;
  }
  clock_gettime(CLOCK_MONOTONIC, &t);
  return (double)(t.tv_sec - t0.tv_sec) + 1e-9 * (double)(t.tv_nsec - t0.tv_nsec);
}

static void write_field(const double *xf, int strf, const double *f, int fcstr, int nfc, int n)
{
  int i, j;
  for (i = 0; i < n; i++) {
    fprintf(stdout, "%1.16e %1.16e %1.16e", xf[(size_t)i * strf + 0], xf[(size_t)i * strf + 1],
            xf[(size_t)i * strf + 2]);
    for (j = 0; j < nfc; j++) fprintf(stdout, " %1.16e", f[(size_t)i * fcstr + j]);
    fprintf(stdout, "\n");
  }
}

#endif
