/* emmi_oracle.c -- CPU restatement of libEMMI_MGFD's multigrid Maxwell path.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP path.  Only tests/,
 * __graft_entry__.smoke() and bench.py's CPU-baseline leg may load it.
 *
 * Written from the algorithm, function by function:
 *   grid construction   src/find_index.c:9-29, src/create_nugrid.c:13-36,
 *                       src/regrid_init_free.c:21-232, src/homogenization.c:15-111
 *   solver              src/gmg.c:39-1932
 *   sources/receivers   src/inject_extract.c:26-785
 *   gradient            src/build_gradient.c:13-111
 * The solver works on one generic axis triple (a, b, c) = (a, a+1, a+2) mod 3
 * so the three edge directions share one code path; the curl-curl operator
 * is invariant under that cyclic relabelling (src/gmg.c:1365-1405).
 * Float/double evaluation order is kept where the reference's results are a
 * bit-exact contract (grid, indices, weights, injection, extraction).
 */
#include "emmi_oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define EMO_PI 3.14159265358979323846264
#define EMO_MU0 (4. * EMO_PI * 1e-7)

static void *xcalloc(size_t n, size_t s)
{
  void *p = calloc(n ? n : 1, s);
  if (!p) { fprintf(stderr, "emmi_oracle: out of memory\n"); exit(1); }
  return p;
}

/* ======================================================================= */
/* grid construction                                                        */
/* ======================================================================= */

/* bisection: largest k with x[k] <= val, clipped to [0, n-2]
 * (src/find_index.c:9-29) */
int emo_find_index(int n, const float *x, float val)
{
  int lo = 0, hi = n - 1, mid;
  while (lo < hi) {
    mid = (lo + hi) / 2;
    if (x[mid] <= val) lo = mid;
    if (x[mid] > val) hi = mid;
    if (hi - lo <= 1) break;
  }
  return lo;
}

/* bisection over x[0..len-1] that reproduces a call with a nominal count n
 * that may exceed len (src/regrid_init_free.c:194); aborts if it would read
 * past len */
static int find_index_guarded(int n, const float *x, int len, float val)
{
  int lo = 0, hi = n - 1, mid;
  while (lo < hi) {
    mid = (lo + hi) / 2;
    if (mid >= len) { fprintf(stderr, "emmi_oracle: find_index reads past the array\n"); exit(1); }
    if (x[mid] <= val) lo = mid;
    if (x[mid] > val) hi = mid;
    if (hi - lo <= 1) break;
  }
  return lo;
}

/* geometric progression x[i] = (q^i - 1) dx/(q - 1) covering len, q by
 * fixed-point iteration in float (src/create_nugrid.c:13-36) */
float emo_create_nugrid(int n, float len, float dx, float *x)
{
  float q = 1.1f, qn;
  float eps = 1e-15;
  int i;
  if (n * dx >= len) {
    for (i = 0; i <= n; i++) x[i] = i * dx;
    return 1;
  }
  for (;;) {
    qn = pow(len * (q - 1.) / dx + 1., 1. / n);
    if (fabs(qn - q) < eps) break;
    q = qn;
  }
  x[0] = 0;
  for (i = 1; i <= n; i++) x[i] = (pow(q, i) - 1.) * dx / (q - 1.);
  return q;
}

/* staggered nodes and widths in float (src/regrid_init_free.c:21-34) */
void emo_staggered(int n, const float *x, float *xs, float *d, float *ds)
{
  int i;
  for (i = 0; i < n; i++) {
    ds[i] = x[i + 1] - x[i];
    xs[i] = 0.5 * (x[i] + x[i + 1]);
  }
  xs[n] = x[n] + 0.5 * ds[n - 1];
  ds[n] = xs[n] - x[n];
  d[0] = xs[0] - x[0];
  for (i = 1; i <= n; i++) d[i] = xs[i] - xs[i - 1];
}

/* 1D overlap average of `in` along a strided axis (src/homogenization.c) */
static void average_axis(const float *in, float *out, int nin, const float *xin,
                         int nout, const float *xout, int nlines,
                         long in_line_stride, long in_step, long out_line_stride, long out_step,
                         const long *line_off_in, const long *line_off_out)
{
  int l, j, i, k;
  for (l = 0; l < nlines; l++) {
    const float *src = in + line_off_in[l];
    float *dst = out + line_off_out[l];
    i = 0;
    for (j = 0; j < nout; j++) {
      float a = xout[j], b = xout[j + 1], s = 0.0;
      while (i < nin && xin[i + 1] <= a) i++;
      for (k = i; k < nin && xin[k] < b; k++) {
        float lo = (xin[k] > a) ? xin[k] : a;
        float hi = (xin[k + 1] < b) ? xin[k + 1] : b;
        float ov = hi - lo;
        if (ov > 0.0) s += src[k * in_step] * ov;
      }
      dst[j * out_step] = s / (b - a);
    }
  }
  (void)in_line_stride; (void)out_line_stride;
}

void emo_homogenize(const float *in, float *out, int nx, int ny, int nz,
                    const float *xx, const float *yy, const float *zz,
                    int n1, int n2, int n3,
                    const float *x1, const float *x2, const float *x3)
{
  float *t1 = xcalloc((size_t)n1 * ny * nz, sizeof(float));
  float *t2 = xcalloc((size_t)n1 * n2 * nz, sizeof(float));
  long *oi, *oo;
  int iy, iz, ix, l;
  /* pass 1: x, lines (iy, iz) in that nesting */
  oi = xcalloc((size_t)ny * nz, sizeof(long)); oo = xcalloc((size_t)ny * nz, sizeof(long));
  for (iz = 0, l = 0; iz < nz; iz++)
    for (iy = 0; iy < ny; iy++, l++) { oi[l] = (long)nx * (iy + (long)ny * iz); oo[l] = (long)n1 * (iy + (long)ny * iz); }
  average_axis(in, t1, nx, xx, n1, x1, ny * nz, 0, 1, 0, 1, oi, oo);
  free(oi); free(oo);
  /* pass 2: y, lines (ix, iz) */
  oi = xcalloc((size_t)n1 * nz, sizeof(long)); oo = xcalloc((size_t)n1 * nz, sizeof(long));
  for (iz = 0, l = 0; iz < nz; iz++)
    for (ix = 0; ix < n1; ix++, l++) { oi[l] = ix + (long)n1 * ny * iz; oo[l] = ix + (long)n1 * n2 * iz; }
  average_axis(t1, t2, ny, yy, n2, x2, n1 * nz, 0, n1, 0, n1, oi, oo);
  free(oi); free(oo);
  /* pass 3: z, lines (ix, iy) */
  oi = xcalloc((size_t)n1 * n2, sizeof(long)); oo = xcalloc((size_t)n1 * n2, sizeof(long));
  for (iy = 0, l = 0; iy < n2; iy++)
    for (ix = 0; ix < n1; ix++, l++) { oi[l] = ix + (long)n1 * iy; oo[l] = ix + (long)n1 * iy; }
  average_axis(t2, out, nz, zz, n3, x3, n1 * n2, 0, (long)n1 * n2, 0, (long)n1 * n2, oi, oo);
  free(oi); free(oo);
  free(t1); free(t2);
}

/* padding, source-centred stretching and homogenisation
 * (src/regrid_init_free.c:37-232) */
void emo_regrid(const emo_model_t *m, float src_x, float src_y, float src_z, emo_grid_t *g)
{
  int nb = m->nb, nxp = m->nx + 2 * nb, nyp = m->ny + 2 * nb, nzp = m->nz + 2 * nb;
  int i, j, k, np1, np2;
  float *x, dist;
  float *s11, *s22, *r33;
  float x1min, x1max, x2min, x2max, x3max;
  size_t np = (size_t)nxp * nyp * nzp;

  for (i = 0; i <= m->nx; i++) g->xx[i + nb] = m->ox + i * m->dx;
  for (i = 0; i <= m->ny; i++) g->yy[i + nb] = m->oy + i * m->dy;
  for (i = 0; i <= m->nz; i++) g->zz[i + nb] = m->oz + i * m->dz;
  x = xcalloc(nb + 1, sizeof(float));
  emo_create_nugrid(nb, m->lextend, m->dx, x);
  for (i = 0; i <= nb; i++) { g->xx[i] = m->ox - x[nb - i]; g->xx[i + m->nx + nb] = m->ox + m->nx * m->dx + x[i]; }
  emo_create_nugrid(nb, m->lextend, m->dy, x);
  for (i = 0; i <= nb; i++) { g->yy[i] = m->oy - x[nb - i]; g->yy[i + m->ny + nb] = m->oy + m->ny * m->dy + x[i]; }
  emo_create_nugrid(nb, m->lextend, m->dz, x);
  for (i = 0; i <= nb; i++) { g->zz[i] = m->oz - x[nb - i]; g->zz[i + m->nz + nb] = m->oz + m->nz * m->dz + x[i]; }
  free(x);
  x1min = g->xx[0]; x1max = g->xx[nxp];
  x2min = g->yy[0]; x2max = g->yy[nyp];
  x3max = g->zz[nzp];

  /* padded model: conductivity for the horizontal terms, resistivity for the
   * vertical one (averaged as a series resistance) */
  s11 = xcalloc(np, sizeof(float)); s22 = xcalloc(np, sizeof(float)); r33 = xcalloc(np, sizeof(float));
#define PIDX(i, j, k) ((i) + (size_t)nxp * ((j) + (size_t)nyp * (k)))
  for (k = 0; k < m->nz; k++)
    for (j = 0; j < m->ny; j++)
      for (i = 0; i < m->nx; i++) {
        size_t s = i + (size_t)m->nx * (j + (size_t)m->ny * k);
        s11[PIDX(i + nb, j + nb, k + nb)] = 1. / m->rho11[s];
        s22[PIDX(i + nb, j + nb, k + nb)] = 1. / m->rho22[s];
        r33[PIDX(i + nb, j + nb, k + nb)] = m->rho33[s];
      }
  for (k = 0; k < nzp; k++)
    for (j = 0; j < nyp; j++)
      for (i = 0; i < nb; i++) {
        s11[PIDX(i, j, k)] = s11[PIDX(nb, j, k)];
        s11[PIDX(nxp - 1 - i, j, k)] = s11[PIDX(nxp - 1 - nb, j, k)];
        s22[PIDX(i, j, k)] = s22[PIDX(nb, j, k)];
        s22[PIDX(nxp - 1 - i, j, k)] = s22[PIDX(nxp - 1 - nb, j, k)];
        r33[PIDX(i, j, k)] = r33[PIDX(nb, j, k)];
        r33[PIDX(nxp - 1 - i, j, k)] = r33[PIDX(nxp - 1 - nb, j, k)];
      }
  for (k = 0; k < nzp; k++)
    for (j = 0; j < nb; j++)
      for (i = 0; i < nxp; i++) {
        s11[PIDX(i, j, k)] = s11[PIDX(i, nb, k)];
        s11[PIDX(i, nyp - 1 - j, k)] = s11[PIDX(i, nyp - 1 - nb, k)];
        s22[PIDX(i, j, k)] = s22[PIDX(i, nb, k)];
        s22[PIDX(i, nyp - 1 - j, k)] = s22[PIDX(i, nyp - 1 - nb, k)];
        r33[PIDX(i, j, k)] = r33[PIDX(i, nb, k)];
        r33[PIDX(i, nyp - 1 - j, k)] = r33[PIDX(i, nyp - 1 - nb, k)];
      }
  for (k = 0; k < nb; k++)
    for (j = 0; j < nyp; j++)
      for (i = 0; i < nxp; i++) {
        s11[PIDX(i, j, k)] = 1. / m->rho_air;
        s11[PIDX(i, j, nzp - 1 - k)] = s11[PIDX(i, j, nzp - 1 - nb)];
        s22[PIDX(i, j, k)] = 1. / m->rho_air;
        s22[PIDX(i, j, nzp - 1 - k)] = s22[PIDX(i, j, nzp - 1 - nb)];
        r33[PIDX(i, j, k)] = m->rho_air;
        r33[PIDX(i, j, nzp - 1 - k)] = r33[PIDX(i, j, nzp - 1 - nb)];
      }
#undef PIDX

  /* x and y: two geometric progressions from the source outwards */
  np1 = m->n1 / 2; np2 = m->n1 - np1;
  x = xcalloc(m->n1 + m->n2 + m->n3 + 3, sizeof(float));
  dist = src_x - x1min;
  emo_create_nugrid(np1, dist, m->d1min, x);
  for (i = 0; i <= np1; i++) g->x1[np1 - i] = src_x - x[i];
  dist = x1max - src_x;
  emo_create_nugrid(np2, dist, m->d1min, x);
  for (i = 0; i <= np2; i++) g->x1[np1 + i] = src_x + x[i];
  np1 = m->n2 / 2; np2 = m->n2 - np1;
  dist = src_y - x2min;
  emo_create_nugrid(np1, dist, m->d2min, x);
  for (i = 0; i <= np1; i++) g->x2[np1 - i] = src_y - x[i];
  dist = x2max - src_y;
  emo_create_nugrid(np2, dist, m->d2min, x);
  for (i = 0; i <= np2; i++) g->x2[np1 + i] = src_y + x[i];
  /* z: input nodes down to 5 cells below the source cell, then a
   * progression; the reference bisects zz with the x count nx+2nb */
  np1 = find_index_guarded(m->nx + 2 * nb, g->zz, nzp + 1, src_z) + 5;
  np2 = m->n3 - np1;
  if (np1 > nzp || np2 < 1) { fprintf(stderr, "emmi_oracle: n3 too small for the source depth\n"); exit(1); }
  for (i = 0; i <= np1; i++) g->x3[i] = g->zz[i];
  dist = x3max - g->x3[np1];
  emo_create_nugrid(np2, dist, m->d3min, x);
  for (i = 0; i <= np2; i++) g->x3[np1 + i] = g->x3[np1] + x[i];
  free(x);

  emo_staggered(m->n1, g->x1, g->x1s, g->d1, g->d1s);
  emo_staggered(m->n2, g->x2, g->x2s, g->d2, g->d2s);
  emo_staggered(m->n3, g->x3, g->x3s, g->d3, g->d3s);

  emo_homogenize(s11, g->sigma11, nxp, nyp, nzp, g->xx, g->yy, g->zz, m->n1, m->n2, m->n3, g->x1, g->x2, g->x3);
  emo_homogenize(s22, g->sigma22, nxp, nyp, nzp, g->xx, g->yy, g->zz, m->n1, m->n2, m->n3, g->x1, g->x2, g->x3);
  emo_homogenize(r33, g->sigma33, nxp, nyp, nzp, g->xx, g->yy, g->zz, m->n1, m->n2, m->n3, g->x1, g->x2, g->x3);
  for (i = 0; i < m->n1 * m->n2 * m->n3; i++) g->sigma33[i] = 1. / g->sigma33[i];
  free(s11); free(s22); free(r33);
}

/* ======================================================================= */
/* multigrid levels                                                         */
/* ======================================================================= */

struct emo_level {
  int n[3];          /* cells per axis */
  int sc[3];         /* coarsening factor relative to the finer level */
  double *x[3], *xs[3], *d[3], *ds[3];
  double *sig[3];    /* sigma_aa * cell volume */
  double *mu;        /* 1/mu_r * cell volume */
};

void emo_level_dims(const emo_level_t *l, int *n123, int *sc)
{
  int a;
  for (a = 0; a < 3; a++) { if (n123) n123[a] = l->n[a]; if (sc) sc[a] = l->sc[a]; }
}
const double *emo_level_cells(const emo_level_t *l, int which) { return which < 3 ? l->sig[which] : l->mu; }
const double *emo_level_axis(const emo_level_t *l, int axis, int kind)
{
  switch (kind) { case 0: return l->x[axis]; case 1: return l->xs[axis]; case 2: return l->d[axis]; default: return l->ds[axis]; }
}

/* src/gmg.c:1672-1685, in double */
static void xs_dx(int n, const double *x, double *xs, double *d, double *ds)
{
  int i;
  for (i = 0; i < n; i++) { ds[i] = x[i + 1] - x[i]; xs[i] = 0.5 * (x[i] + x[i + 1]); }
  xs[n] = x[n] + 0.5 * ds[n - 1];
  ds[n] = xs[n] - x[n];
  for (i = 1; i <= n; i++) d[i] = xs[i] - xs[i - 1];
  d[0] = xs[0] - x[0];
}

static emo_level_t *level_alloc(const int *n)
{
  emo_level_t *l = xcalloc(1, sizeof(*l));
  size_t nc = (size_t)n[0] * n[1] * n[2];
  int a;
  for (a = 0; a < 3; a++) {
    l->n[a] = n[a];
    l->sc[a] = 1;
    l->x[a] = xcalloc(n[a] + 1, sizeof(double));
    l->xs[a] = xcalloc(n[a] + 1, sizeof(double));
    l->d[a] = xcalloc(n[a] + 1, sizeof(double));
    l->ds[a] = xcalloc(n[a] + 1, sizeof(double));
    l->sig[a] = xcalloc(nc, sizeof(double));
  }
  l->mu = xcalloc(nc, sizeof(double));
  return l;
}

void emo_level_free(emo_level_t *l)
{
  int a;
  if (!l) return;
  for (a = 0; a < 3; a++) { free(l->x[a]); free(l->xs[a]); free(l->d[a]); free(l->ds[a]); free(l->sig[a]); }
  free(l->mu);
  free(l);
}

/* level 0 of src/gmg.c:1736-1802 */
emo_level_t *emo_level0_new(int n1, int n2, int n3,
                            const float *x1, const float *x2, const float *x3,
                            const float *sig11, const float *sig22, const float *sig33)
{
  int n[3] = {n1, n2, n3}, a, i, j, k;
  const float *xf[3] = {x1, x2, x3};
  emo_level_t *l = level_alloc(n);
  for (a = 0; a < 3; a++) {
    for (i = 0; i <= n[a]; i++) l->x[a][i] = xf[a][i];
    xs_dx(n[a], l->x[a], l->xs[a], l->d[a], l->ds[a]);
  }
  for (k = 0; k < n3; k++)
    for (j = 0; j < n2; j++)
      for (i = 0; i < n1; i++) {
        size_t c = i + (size_t)n1 * (j + (size_t)n2 * k);
        double vol = l->ds[0][i] * l->ds[1][j] * l->ds[2][k];
        l->sig[0][c] = sig11[c] * vol;
        l->sig[1][c] = sig22[c] * vol;
        l->sig[2][c] = sig33[c] * vol;
        l->mu[c] = 1.0f * vol;
      }
  return l;
}

/* coarse level of src/gmg.c:1744-1858: every sc-th node, medium summed over
 * the (repeated) eight children and divided by the repeat count */
emo_level_t *emo_level_coarsen(const emo_level_t *f, int icycle, int isemi)
{
  int sc[3], n[3], a, i, j, k, repeat, q;
  emo_level_t *l;
  for (a = 0; a < 3; a++) sc[a] = (f->n[a] % 2 == 0 && f->n[a] > 2) ? 2 : 1;
  if (isemi) sc[icycle % 3] = 1;
  for (a = 0; a < 3; a++) n[a] = f->n[a] / sc[a];
  l = level_alloc(n);
  for (a = 0; a < 3; a++) {
    l->sc[a] = sc[a];
    for (i = 0; i <= n[a]; i++) l->x[a][i] = f->x[a][sc[a] * i];
    xs_dx(n[a], l->x[a], l->xs[a], l->d[a], l->ds[a]);
  }
  repeat = 8 / (sc[0] * sc[1] * sc[2]);
  for (k = 0; k < n[2]; k++) {
    int k0 = sc[2] == 2 ? 2 * k : k, k1 = sc[2] == 2 ? 2 * k + 1 : k;
    for (j = 0; j < n[1]; j++) {
      int j0 = sc[1] == 2 ? 2 * j : j, j1 = sc[1] == 2 ? 2 * j + 1 : j;
      for (i = 0; i < n[0]; i++) {
        int i0 = sc[0] == 2 ? 2 * i : i, i1 = sc[0] == 2 ? 2 * i + 1 : i;
        size_t c = i + (size_t)n[0] * (j + (size_t)n[1] * k);
#define FC(ii, jj, kk) ((ii) + (size_t)f->n[0] * ((jj) + (size_t)f->n[1] * (kk)))
        for (q = 0; q < 4; q++) {
          const double *s = q < 3 ? f->sig[q] : f->mu;
          double v = s[FC(i0, j0, k0)] + s[FC(i1, j0, k0)] + s[FC(i0, j1, k0)] + s[FC(i1, j1, k0)]
                   + s[FC(i0, j0, k1)] + s[FC(i1, j0, k1)] + s[FC(i0, j1, k1)] + s[FC(i1, j1, k1)];
          if (q < 3) l->sig[q][c] = v / repeat; else l->mu[c] = v / repeat;
        }
#undef FC
      }
    }
  }
  return l;
}

/* ---- indexing helpers ---------------------------------------------------- */
typedef struct {
  int n[3];
  long es[3];   /* edge-array strides */
  long ncomp;   /* entries per component */
  long cs[3];   /* cell strides */
} geom_t;

static geom_t geom_of(const emo_level_t *l)
{
  geom_t g;
  g.n[0] = l->n[0]; g.n[1] = l->n[1]; g.n[2] = l->n[2];
  g.es[0] = 1; g.es[1] = l->n[0] + 1; g.es[2] = (long)(l->n[0] + 1) * (l->n[1] + 1);
  g.ncomp = g.es[2] * (l->n[2] + 1);
  g.cs[0] = 1; g.cs[1] = l->n[0]; g.cs[2] = (long)l->n[0] * l->n[1];
  return g;
}

static long eidx(const geom_t *g, int comp, const int *p)
{
  return comp * g->ncomp + p[0] * g->es[0] + p[1] * g->es[1] + p[2] * g->es[2];
}
static long cidx(const geom_t *g, const int *p)
{
  return p[0] * g->cs[0] + p[1] * g->cs[1] + p[2] * g->cs[2];
}

/* interior edge of component `a` at p: p_a < n_a, 0 < p_b < n_b, 0 < p_c < n_c */
static int interior(const geom_t *g, int a, const int *p)
{
  int b = (a + 1) % 3, c = (a + 2) % 3;
  return p[a] >= 0 && p[a] < g->n[a] && p[b] > 0 && p[b] < g->n[b] && p[c] > 0 && p[c] < g->n[c];
}

/* one row of the curl-curl operator: the 13 couplings of E_a at p
 * (src/gmg.c:1365-1405 rewritten with explicit coefficients) */
typedef struct {
  int comp[13];
  int pos[13][3];
  ocplx coef[13];
} row_t;

static void stencil_row(const emo_level_t *l, const geom_t *g, double omu, int a, const int *p, row_t *r)
{
  int b = (a + 1) % 3, c = (a + 2) % 3, t, q[3];
  double dsa = l->ds[a][p[a]], dsb = l->ds[b][p[b]], dsbm = l->ds[b][p[b] - 1];
  double dsc = l->ds[c][p[c]], dscm = l->ds[c][p[c] - 1];
  double m1, m2, m3, m4, sbar;
  int pb[3], pc[3], pbc[3];
  ocplx iom = CMPLX(0.0, omu);
  memcpy(pb, p, sizeof pb); pb[b]--;
  memcpy(pc, p, sizeof pc); pc[c]--;
  memcpy(pbc, pb, sizeof pbc); pbc[c]--;
  m1 = 0.5 * (l->mu[cidx(g, p)] + l->mu[cidx(g, pc)]);
  m2 = 0.5 * (l->mu[cidx(g, pb)] + l->mu[cidx(g, pbc)]);
  m3 = 0.5 * (l->mu[cidx(g, p)] + l->mu[cidx(g, pb)]);
  m4 = 0.5 * (l->mu[cidx(g, pc)] + l->mu[cidx(g, pbc)]);
  sbar = (b < c) ? 0.25 * (l->sig[a][cidx(g, p)] + l->sig[a][cidx(g, pb)] + l->sig[a][cidx(g, pc)] + l->sig[a][cidx(g, pbc)])
                 : 0.25 * (l->sig[a][cidx(g, p)] + l->sig[a][cidx(g, pc)] + l->sig[a][cidx(g, pb)] + l->sig[a][cidx(g, pbc)]);
  t = 0;
#define PUT(cmp, d0, s0, d1, s1, val) do { \
    memcpy(q, p, sizeof q); q[d0] += s0; q[d1] += s1; \
    r->comp[t] = cmp; memcpy(r->pos[t], q, sizeof q); r->coef[t] = (val); t++; } while (0)
  PUT(a, a, 0, a, 0, (m1 / (dsb * dsb) + m2 / (dsbm * dsbm)) + (m3 / (dsc * dsc) + m4 / (dscm * dscm)) - iom * sbar);
  PUT(a, b, +1, a, 0, -m1 / (dsb * dsb));
  PUT(a, b, -1, a, 0, -m2 / (dsbm * dsbm));
  PUT(a, c, +1, a, 0, -m3 / (dsc * dsc));
  PUT(a, c, -1, a, 0, -m4 / (dscm * dscm));
  PUT(b, a, +1, a, 0, m1 / (dsa * dsb));
  PUT(b, a, 0, a, 0, -m1 / (dsa * dsb));
  PUT(b, b, -1, a, +1, -m2 / (dsa * dsbm));
  PUT(b, b, -1, a, 0, m2 / (dsa * dsbm));
  PUT(c, a, +1, a, 0, m3 / (dsa * dsc));
  PUT(c, a, 0, a, 0, -m3 / (dsa * dsc));
  PUT(c, c, -1, a, +1, -m4 / (dsa * dscm));
  PUT(c, c, -1, a, 0, m4 / (dsa * dscm));
#undef PUT
}

/* (A u) on row (a, p) in the reference's curl-first evaluation order */
static ocplx apply_row(const emo_level_t *l, const geom_t *g, double omu, int a, const int *p, const ocplx *u)
{
  int b = (a + 1) % 3, c = (a + 2) % 3;
  int q[3], pb[3], pc[3], pbc[3];
  double dsa = l->ds[a][p[a]], dsb = l->ds[b][p[b]], dsbm = l->ds[b][p[b] - 1];
  double dsc = l->ds[c][p[c]], dscm = l->ds[c][p[c] - 1];
  double m1, m2, m3, m4, sbar;
  ocplx t1, t2, t3, t4, ea;
  ocplx iom = CMPLX(0.0, omu);
  long o = eidx(g, 0, p), sa = g->es[a], sb = g->es[b], sc = g->es[c];
  const ocplx *ua = u + a * g->ncomp + o, *ub = u + b * g->ncomp + o, *uc = u + c * g->ncomp + o;
  (void)q;
  memcpy(pb, p, sizeof pb); pb[b]--;
  memcpy(pc, p, sizeof pc); pc[c]--;
  memcpy(pbc, pb, sizeof pbc); pbc[c]--;
  ea = ua[0];
  /* curl along c on the face above (t1) and below (t2) in b */
  t1 = (ub[sa] - ub[0]) / dsa - (ua[sb] - ea) / dsb;
  t2 = (ub[sa - sb] - ub[-sb]) / dsa - (ea - ua[-sb]) / dsbm;
  /* curl along b on the faces above (t3) and below (t4) in c */
  t3 = (ua[sc] - ea) / dsc - (uc[sa] - uc[0]) / dsa;
  t4 = (ea - ua[-sc]) / dscm - (uc[sa - sc] - uc[-sc]) / dsa;
  m1 = 0.5 * (l->mu[cidx(g, p)] + l->mu[cidx(g, pc)]);
  m2 = 0.5 * (l->mu[cidx(g, pb)] + l->mu[cidx(g, pbc)]);
  m3 = 0.5 * (l->mu[cidx(g, p)] + l->mu[cidx(g, pb)]);
  m4 = 0.5 * (l->mu[cidx(g, pc)] + l->mu[cidx(g, pbc)]);
  sbar = (b < c) ? 0.25 * (l->sig[a][cidx(g, p)] + l->sig[a][cidx(g, pb)] + l->sig[a][cidx(g, pc)] + l->sig[a][cidx(g, pbc)])
                 : 0.25 * (l->sig[a][cidx(g, p)] + l->sig[a][cidx(g, pc)] + l->sig[a][cidx(g, pb)] + l->sig[a][cidx(g, pbc)]);
  t1 *= m1; t2 *= m2; t3 *= m3; t4 *= m4;
  return ((t1 / dsb - t2 / dsbm) - (t3 / dsc - t4 / dscm)) - iom * sbar * ea;
}

/* src/gmg.c:1316-1410 */
void emo_residual(const emo_level_t *l, double omu, const ocplx *u, const ocplx *f, ocplx *r)
{
  geom_t g = geom_of(l);
  int a, p[3];
  memcpy(r, f, sizeof(ocplx) * 3 * g.ncomp);
  for (p[2] = 0; p[2] < g.n[2]; p[2]++)
    for (p[1] = 0; p[1] < g.n[1]; p[1]++)
      for (p[0] = 0; p[0] < g.n[0]; p[0]++)
        for (a = 0; a < 3; a++)
          if (interior(&g, a, p)) r[eidx(&g, a, p)] -= apply_row(l, &g, omu, a, p, u);
}

void emo_apply(const emo_level_t *l, double omu, const ocplx *u, ocplx *y)
{
  geom_t g = geom_of(l);
  int a, p[3];
  memset(y, 0, sizeof(ocplx) * 3 * g.ncomp);
  for (p[2] = 0; p[2] < g.n[2]; p[2]++)
    for (p[1] = 0; p[1] < g.n[1]; p[1]++)
      for (p[0] = 0; p[0] < g.n[0]; p[0]++)
        for (a = 0; a < 3; a++)
          if (interior(&g, a, p)) y[eidx(&g, a, p)] = apply_row(l, &g, omu, a, p, u);
}

/* ---- block solves ----------------------------------------------------------- */

/* dense LU without pivoting, in place, then substitution (src/gmg.c:39-66) */
static void dense_solve(ocplx *A, ocplx *x, int n)
{
  int i, j, k;
  for (k = 0; k < n - 1; k++) {
    for (i = k + 1; i < n; i++) A[i * n + k] /= A[k * n + k];
    for (i = k + 1; i < n; i++)
      for (j = k + 1; j < n; j++) A[i * n + j] -= A[i * n + k] * A[k * n + j];
  }
  for (i = 0; i < n; i++) { ocplx s = 0; for (j = 0; j < i; j++) s += A[i * n + j] * x[j]; x[i] -= s; }
  for (i = n - 1; i >= 0; i--) { ocplx s = 0; for (j = i + 1; j < n; j++) s += A[i * n + j] * x[j]; x[i] = (x[i] - s) / A[i * n + i]; }
}

/* banded LU without pivoting on a matrix of half-bandwidth 5 stored by
 * diagonals, A[r*11 + (col - r + 5)]; fill-in of a banded LU stays inside the
 * band, so this is an exact solve (src/gmg.c:71-112) */
#define BW 5
static void band_solve(ocplx *A, ocplx *x, int n)
{
  int r, i, j;
  for (r = 0; r < n - 1; r++) {
    int last = r + BW < n - 1 ? r + BW : n - 1;
    for (i = r + 1; i <= last; i++) A[i * 11 + (r - i + BW)] /= A[r * 11 + BW];
    for (i = r + 1; i <= last; i++)
      for (j = r + 1; j <= last; j++) A[i * 11 + (j - i + BW)] -= A[i * 11 + (r - i + BW)] * A[r * 11 + (j - r + BW)];
  }
  for (i = 0; i < n; i++) {
    ocplx s = 0;
    for (j = (i - BW > 0 ? i - BW : 0); j < i; j++) s += A[i * 11 + (j - i + BW)] * x[j];
    x[i] -= s;
  }
  for (i = n - 1; i >= 0; i--) {
    ocplx s = 0;
    for (j = i + 1; j < n && j <= i + BW; j++) s += A[i * 11 + (j - i + BW)] * x[j];
    x[i] = (x[i] - s) / A[i * 11 + BW];
  }
}

/* Position of edge (comp, q) in the unknowns of line (a; pb, pc), or -1.
 * Per node i = 1..n_a-1 the five unknowns are E_a(i-1), then the two edges
 * of the lower-numbered perpendicular axis lo (positions p_lo-1, p_lo), then
 * the two of the higher one hi; E_a(n_a-1) closes the line.  This is the
 * unknown order of src/gmg.c:414-637 (x), :736-955 (y), :1056-1275 (z). */
static int line_row(const geom_t *g, int a, int pb, int pc, int comp, const int *q)
{
  int b = (a + 1) % 3, c = (a + 2) % 3, na = g->n[a];
  int lo = b < c ? b : c, hi = b < c ? c : b;
  int p[3];
  p[b] = pb; p[c] = pc;
  if (comp == a) {
    if (q[b] == pb && q[c] == pc && q[a] >= 0 && q[a] <= na - 1) return 5 * q[a];
    return -1;
  }
  if (q[a] < 1 || q[a] > na - 1) return -1;
  if (comp == lo) {
    if (q[hi] == p[hi] && (q[lo] == p[lo] - 1 || q[lo] == p[lo])) return 5 * (q[a] - 1) + 1 + (q[lo] - (p[lo] - 1));
    return -1;
  }
  if (q[lo] == p[lo] && (q[hi] == p[hi] - 1 || q[hi] == p[hi])) return 5 * (q[a] - 1) + 3 + (q[hi] - (p[hi] - 1));
  return -1;
}

static void line_solve(const emo_level_t *l, const geom_t *g, double omu, int a, int pb, int pc,
                       ocplx *u, const ocplx *f, ocplx *A, ocplx *rhs, int (*rowedge)[4])
{
  int b = (a + 1) % 3, c = (a + 2) % 3, na = g->n[a], nr = 5 * na - 4, r, t, i;
  int lo = b < c ? b : c, hi = b < c ? c : b;
  row_t st;
  /* enumerate the line's edges: rowedge[r] = (comp, q0, q1, q2) */
  for (i = 1; i <= na - 1; i++) {
    int base = 5 * (i - 1), s;
    for (s = 0; s < 5; s++) {
      int q[3];
      q[a] = i; q[b] = pb; q[c] = pc;
      if (s == 0) { rowedge[base][0] = a; q[a] = i - 1; }
      else if (s <= 2) { rowedge[base + s][0] = lo; q[lo] += s - 2; }
      else { rowedge[base + s][0] = hi; q[hi] += s - 4; }
      rowedge[base + s][1] = q[0]; rowedge[base + s][2] = q[1]; rowedge[base + s][3] = q[2];
    }
  }
  rowedge[nr - 1][0] = a;
  { int q[3]; q[a] = na - 1; q[b] = pb; q[c] = pc; rowedge[nr - 1][1] = q[0]; rowedge[nr - 1][2] = q[1]; rowedge[nr - 1][3] = q[2]; }
  /* zero the line's unknowns (src/gmg.c:392-403) */
  for (r = 0; r < nr; r++) u[eidx(g, rowedge[r][0], &rowedge[r][1])] = 0;
  memset(A, 0, sizeof(ocplx) * 11 * nr);
  for (r = 0; r < nr; r++) {
    int comp = rowedge[r][0];
    const int *q = &rowedge[r][1];
    ocplx acc;
    stencil_row(l, g, omu, comp, q, &st);
    /* off-line couplings go to the right-hand side (line values are zero) */
    acc = apply_row(l, g, omu, comp, q, u);
    rhs[r] = f[eidx(g, comp, q)] - acc;
    for (t = 0; t < 13; t++) {
      int col = line_row(g, a, pb, pc, st.comp[t], st.pos[t]);
      if (col < 0) continue;
      if (col - r > BW || r - col > BW) { fprintf(stderr, "emmi_oracle: band overflow\n"); exit(1); }
      A[r * 11 + (col - r + BW)] += st.coef[t];
    }
  }
  band_solve(A, rhs, nr);
  for (r = 0; r < nr; r++) u[eidx(g, rowedge[r][0], &rowedge[r][1])] = rhs[r];
}

void emo_line_solve(const emo_level_t *l, double omu, int a, int pb, int pc, ocplx *u, const ocplx *f)
{
  geom_t g = geom_of(l);
  int nr = 5 * g.n[a] - 4;
  ocplx *A = xcalloc((size_t)11 * nr, sizeof(ocplx)), *rhs = xcalloc(nr, sizeof(ocplx));
  int (*re)[4] = xcalloc(nr, sizeof(*re));
  line_solve(l, &g, omu, a, pb, pc, u, f, A, rhs, re);
  free(A); free(rhs); free(re);
}

/* src/gmg.c:340-1298: lines ordered with the higher-numbered perpendicular
 * axis outermost, reversed when iter is odd */
void emo_line_gs(const emo_level_t *l, double omu, int a, int iter, int colour_order, ocplx *u, const ocplx *f)
{
  geom_t g = geom_of(l);
  int b = (a + 1) % 3, c = (a + 2) % 3, nr = 5 * g.n[a] - 4;
  int outer = b > c ? b : c, inner = b > c ? c : b;
  int no = g.n[outer] - 1, ni = g.n[inner] - 1, io, ii, colour, nc = colour_order ? 4 : 1;
  ocplx *A = xcalloc((size_t)11 * nr, sizeof(ocplx)), *rhs = xcalloc(nr, sizeof(ocplx));
  int (*re)[4] = xcalloc(nr, sizeof(*re));
  for (colour = 0; colour < nc; colour++) {
    int col = (iter % 2 == 1) ? nc - 1 - colour : colour;
    for (io = 0; io < no; io++) {
      int vo = (iter % 2 == 1) ? no - io : io + 1;
      for (ii = 0; ii < ni; ii++) {
        int vi = (iter % 2 == 1) ? ni - ii : ii + 1;
        int pos[3], pbv, pcv;
        pos[outer] = vo; pos[inner] = vi;
        pbv = pos[b]; pcv = pos[c];
        if (colour_order && ((pbv & 1) | ((pcv & 1) << 1)) != col) continue;
        line_solve(l, &g, omu, a, pbv, pcv, u, f, A, rhs, re);
      }
    }
  }
  free(A); free(rhs); free(re);
}

/* src/gmg.c:116-336: vertex patch of the six edges around node (i,j,k) */
void emo_point_gs(const emo_level_t *l, double omu, int iter, int colour_order, ocplx *u, const ocplx *f)
{
  geom_t g = geom_of(l);
  int n1 = g.n[0], n2 = g.n[1], n3 = g.n[2], ii, jj, kk, colour, nc = colour_order ? 8 : 1;
  row_t st;
  for (colour = 0; colour < nc; colour++) {
    int col = (iter % 2 == 1) ? nc - 1 - colour : colour;
    for (kk = 0; kk < n3 - 1; kk++)
      for (jj = 0; jj < n2 - 1; jj++)
        for (ii = 0; ii < n1 - 1; ii++) {
          int i = (iter % 2 == 1) ? n1 - 1 - ii : ii + 1;
          int j = (iter % 2 == 1) ? n2 - 1 - jj : jj + 1;
          int k = (iter % 2 == 1) ? n3 - 1 - kk : kk + 1;
          int node[3] = {i, j, k}, e[6][4], s, t, m;
          ocplx A[36], rhs[6];
          if (colour_order && ((i & 1) | ((j & 1) << 1) | ((k & 1) << 2)) != col) continue;
          /* patch order: E_x(i-1), E_x(i), E_y(j-1), E_y(j), E_z(k-1), E_z(k) */
          for (s = 0; s < 6; s++) {
            int cmp = s / 2;
            e[s][0] = cmp; e[s][1] = node[0]; e[s][2] = node[1]; e[s][3] = node[2];
            if (s % 2 == 0) e[s][1 + cmp]--;
            u[eidx(&g, cmp, &e[s][1])] = 0;
          }
          memset(A, 0, sizeof A);
          for (s = 0; s < 6; s++) {
            stencil_row(l, &g, omu, e[s][0], &e[s][1], &st);
            rhs[s] = f[eidx(&g, e[s][0], &e[s][1])] - apply_row(l, &g, omu, e[s][0], &e[s][1], u);
            for (t = 0; t < 13; t++)
              for (m = 0; m < 6; m++)
                if (st.comp[t] == e[m][0] && st.pos[t][0] == e[m][1] && st.pos[t][1] == e[m][2] && st.pos[t][2] == e[m][3])
                  A[s * 6 + m] += st.coef[t];
          }
          dense_solve(A, rhs, 6);
          for (s = 0; s < 6; s++) u[eidx(&g, e[s][0], &e[s][1])] = rhs[s];
        }
  }
}

/* weights of fine node `ff` between coarse nodes (src/gmg.c:1478-1488) */
static void tweights(const emo_level_t *f, const emo_level_t *c, int a, int ff, int *ic, double *wl, double *wr)
{
  int i = c->sc[a] == 2 ? ff / 2 : ff;
  *ic = i;
  *wr = (f->x[a][ff] - c->x[a][i]) / (c->x[a][i + 1] - c->x[a][i]);
  *wl = 1. - *wr;
}

/* src/gmg.c:1510-1573: exact transpose of the prolongation */
void emo_restrict(const emo_level_t *fl, const emo_level_t *cl, const ocplx *rf, ocplx *fc)
{
  geom_t gf = geom_of(fl), gc = geom_of(cl);
  int p[3], ic[3], a;
  double wl[3], wr[3];
  memset(fc, 0, sizeof(ocplx) * 3 * gc.ncomp);
  for (p[2] = 0; p[2] < gf.n[2]; p[2]++)
    for (p[1] = 0; p[1] < gf.n[1]; p[1]++)
      for (p[0] = 0; p[0] < gf.n[0]; p[0]++) {
        for (a = 0; a < 3; a++) tweights(fl, cl, a, p[a], &ic[a], &wl[a], &wr[a]);
        for (a = 0; a < 3; a++) {
          int b = (a + 1) % 3, c = (a + 2) % 3, q[3];
          ocplx v;
          if (!(p[b] > 0 && p[c] > 0)) continue;
          v = rf[eidx(&gf, a, p)];
          /* the two perpendicular axes in increasing index order */
          {
            int lo = b < c ? b : c, hi = b < c ? c : b;
            q[a] = ic[a];
            q[lo] = ic[lo]; q[hi] = ic[hi];         fc[eidx(&gc, a, q)] += wl[lo] * wl[hi] * v;
            q[lo] = ic[lo]; q[hi] = ic[hi] + 1;     fc[eidx(&gc, a, q)] += wl[lo] * wr[hi] * v;
            q[lo] = ic[lo] + 1; q[hi] = ic[hi];     fc[eidx(&gc, a, q)] += wr[lo] * wl[hi] * v;
            q[lo] = ic[lo] + 1; q[hi] = ic[hi] + 1; fc[eidx(&gc, a, q)] += wr[lo] * wr[hi] * v;
          }
        }
      }
}

/* src/gmg.c:1463-1505: constant along the edge, bilinear across it */
void emo_prolong(const emo_level_t *fl, const emo_level_t *cl, const ocplx *uc, ocplx *uf)
{
  geom_t gf = geom_of(fl), gc = geom_of(cl);
  int p[3], ic[3], a;
  double wl[3], wr[3];
  for (p[2] = 0; p[2] < gf.n[2]; p[2]++)
    for (p[1] = 0; p[1] < gf.n[1]; p[1]++)
      for (p[0] = 0; p[0] < gf.n[0]; p[0]++) {
        for (a = 0; a < 3; a++) tweights(fl, cl, a, p[a], &ic[a], &wl[a], &wr[a]);
        for (a = 0; a < 3; a++) {
          int b = (a + 1) % 3, c = (a + 2) % 3, q[3], lo, hi;
          ocplx v00, v01, v10, v11;
          if (!(p[b] > 0 && p[c] > 0)) continue;
          lo = b < c ? b : c; hi = b < c ? c : b;
          q[a] = ic[a];
          q[lo] = ic[lo]; q[hi] = ic[hi];         v00 = uc[eidx(&gc, a, q)];
          q[lo] = ic[lo]; q[hi] = ic[hi] + 1;     v01 = uc[eidx(&gc, a, q)];
          q[lo] = ic[lo] + 1; q[hi] = ic[hi];     v10 = uc[eidx(&gc, a, q)];
          q[lo] = ic[lo] + 1; q[hi] = ic[hi] + 1; v11 = uc[eidx(&gc, a, q)];
          uf[eidx(&gf, a, p)] += wl[lo] * (wl[hi] * v00 + wr[hi] * v01) + wr[lo] * (wl[hi] * v10 + wr[hi] * v11);
        }
      }
}

/* src/gmg.c:1414-1460: H = curl E / (i omega mu0) on every cell-face index */
void emo_curl_h(const emo_level_t *l, double omu, const ocplx *u, ocplx *h)
{
  geom_t g = geom_of(l);
  int p[3], a;
  ocplx iom = CMPLX(0.0, omu);
  memset(h, 0, sizeof(ocplx) * 3 * g.ncomp);
  for (p[2] = 0; p[2] < g.n[2]; p[2]++)
    for (p[1] = 0; p[1] < g.n[1]; p[1]++)
      for (p[0] = 0; p[0] < g.n[0]; p[0]++)
        for (a = 0; a < 3; a++) {
          int b = (a + 1) % 3, c = (a + 2) % 3, qb[3], qc[3];
          ocplx t;
          memcpy(qb, p, sizeof qb); qb[b]++;
          memcpy(qc, p, sizeof qc); qc[c]++;
          /* H_a = d_b E_c - d_c E_b */
          t = (u[eidx(&g, c, qb)] - u[eidx(&g, c, p)]) / l->ds[b][p[b]] - (u[eidx(&g, b, qc)] - u[eidx(&g, b, p)]) / l->ds[c][p[c]];
          h[eidx(&g, a, p)] = t / iom;
        }
}

/* ---- cycles ----------------------------------------------------------------- */
typedef struct {
  emo_params_t p;
  int lmax, icycle, nhist;
  double omu, rnorm0;
  double *hist;
  emo_level_t **lev;
  ocplx **u, **f, **r;
} solver_t;

static int depth(int n)
{
  int l = 0;
  while (n % 2 == 0 && n > 1) { l++; n >>= 1; }
  if (n > 2) l++;
  return l;
}

static long level_len(const emo_level_t *l) { return 3L * (l->n[0] + 1) * (l->n[1] + 1) * (l->n[2] + 1); }

static void smooth(solver_t *s, int lev, int iter)
{
  if (s->p.isemicoarsen && lev < s->lmax - 1) {
    int a;
    for (a = 0; a < 3; a++)
      if (s->lev[lev + 1]->sc[a] == 2) emo_line_gs(s->lev[lev], s->omu, a, iter, s->p.colour_order, s->u[lev], s->f[lev]);
  } else {
    emo_point_gs(s->lev[lev], s->omu, iter, s->p.colour_order, s->u[lev], s->f[lev]);
  }
}

static int coarsens(solver_t *s, int lev)
{
  return lev < s->lmax - 1 && s->lev[lev + 1]->sc[0] * s->lev[lev + 1]->sc[1] * s->lev[lev + 1]->sc[2] > 1;
}

/* level-0 convergence check shared by the three cycle types; returns 1 when
 * converged (src/gmg.c:1594-1600) */
static int check_norm(solver_t *s)
{
  long n = level_len(s->lev[0]), i;
  double acc = 0, rn;
  for (i = 0; i < n; i++) acc += creal(conj(s->r[0][i]) * s->r[0][i]);
  rn = sqrt(acc);
  if (s->hist) s->hist[s->nhist] = rn;
  s->nhist++;
  if (s->p.verb) printf("icycle=%d rnorm=%e\n", s->icycle, rn);
  if (s->icycle == 0) s->rnorm0 = rn;
  else if (rn < s->rnorm0 * s->p.tol) { s->icycle = s->p.ncycle; return 1; }
  return 0;
}

static void v_cycle(solver_t *s, int lev)
{
  int i;
  for (i = 0; i < s->p.v1; i++) smooth(s, lev, i);
  if (coarsens(s, lev)) {
    emo_residual(s->lev[lev], s->omu, s->u[lev], s->f[lev], s->r[lev]);
    if (s->p.cycleopt == 1 && lev == 0 && check_norm(s)) return;
    emo_restrict(s->lev[lev], s->lev[lev + 1], s->r[lev], s->f[lev + 1]);
    memset(s->u[lev + 1], 0, sizeof(ocplx) * level_len(s->lev[lev + 1]));
    v_cycle(s, lev + 1);
    emo_prolong(s->lev[lev], s->lev[lev + 1], s->u[lev + 1], s->u[lev]);
  }
  for (i = 0; i < s->p.v2; i++) smooth(s, lev, i);
}

static void w_cycle(solver_t *s, int lev)
{
  int i;
  for (i = 0; i < s->p.v1; i++) smooth(s, lev, i);
  if (coarsens(s, lev)) {
    emo_residual(s->lev[lev], s->omu, s->u[lev], s->f[lev], s->r[lev]);
    if (s->p.cycleopt == 3 && lev == 0 && check_norm(s)) return;
    emo_restrict(s->lev[lev], s->lev[lev + 1], s->r[lev], s->f[lev + 1]);
    memset(s->u[lev + 1], 0, sizeof(ocplx) * level_len(s->lev[lev + 1]));
    w_cycle(s, lev + 1);
    w_cycle(s, lev + 1);
    emo_prolong(s->lev[lev], s->lev[lev + 1], s->u[lev + 1], s->u[lev]);
  }
  for (i = 0; i < s->p.v2; i++) smooth(s, lev, i);
}

static void f_cycle(solver_t *s, int lev)
{
  if (lev == s->lmax - 1) {
    memset(s->u[lev], 0, sizeof(ocplx) * level_len(s->lev[lev]));
  } else if (coarsens(s, lev)) {
    emo_residual(s->lev[lev], s->omu, s->u[lev], s->f[lev], s->r[lev]);
    if (s->p.cycleopt == 2 && lev == 0 && check_norm(s)) return;
    emo_restrict(s->lev[lev], s->lev[lev + 1], s->r[lev], s->f[lev + 1]);
    memset(s->u[lev + 1], 0, sizeof(ocplx) * level_len(s->lev[lev + 1]));
    f_cycle(s, lev + 1);
    emo_prolong(s->lev[lev], s->lev[lev + 1], s->u[lev + 1], s->u[lev]);
  }
  v_cycle(s, lev);
}

void emo_default_params(emo_params_t *p)
{
  p->cycleopt = 2; p->ncycle = 30; p->v1 = 1; p->v2 = 1; p->isemicoarsen = 1; p->tol = 1e-6;
  p->colour_order = 0; p->verb = 0;
}

/* src/gmg.c:1894-1932 */
int emo_gmg_solve(int n1, int n2, int n3, const float *x1, const float *x2, const float *x3,
                  const float *sig11, const float *sig22, const float *sig33, float omega,
                  const emo_params_t *p, ocplx *b, ocplx *x, double *hist)
{
  solver_t s;
  int lev, i, j, k;
  long n;
  memset(&s, 0, sizeof s);
  s.p = *p;
  s.hist = hist;
  s.omu = (double)omega * EMO_MU0;
  s.lmax = depth(n1);
  if (depth(n2) > s.lmax) s.lmax = depth(n2);
  if (depth(n3) > s.lmax) s.lmax = depth(n3);
  s.lev = xcalloc(s.lmax, sizeof(*s.lev));
  s.u = xcalloc(s.lmax, sizeof(*s.u)); s.f = xcalloc(s.lmax, sizeof(*s.f)); s.r = xcalloc(s.lmax, sizeof(*s.r));
  s.lev[0] = emo_level0_new(n1, n2, n3, x1, x2, x3, sig11, sig22, sig33);
  n = level_len(s.lev[0]);
  s.u[0] = xcalloc(n, sizeof(ocplx)); s.f[0] = xcalloc(n, sizeof(ocplx)); s.r[0] = xcalloc(n, sizeof(ocplx));
  memcpy(s.f[0], b, n * sizeof(ocplx));
  {
    geom_t g = geom_of(s.lev[0]);
    const emo_level_t *l0 = s.lev[0];
    for (k = 0; k <= n3; k++)
      for (j = 0; j <= n2; j++)
        for (i = 0; i <= n1; i++) {
          long e = i + g.es[1] * j + g.es[2] * k;
          s.f[0][e] *= l0->ds[0][i] * l0->d[1][j] * l0->d[2][k];
          s.f[0][g.ncomp + e] *= l0->d[0][i] * l0->ds[1][j] * l0->d[2][k];
          s.f[0][2 * g.ncomp + e] *= l0->d[0][i] * l0->d[1][j] * l0->ds[2][k];
        }
  }
  for (s.icycle = 0; s.icycle < s.p.ncycle; s.icycle++) {
    for (lev = 1; lev < s.lmax; lev++) {
      long nl;
      s.lev[lev] = emo_level_coarsen(s.lev[lev - 1], s.icycle, s.p.isemicoarsen);
      nl = level_len(s.lev[lev]);
      s.u[lev] = xcalloc(nl, sizeof(ocplx)); s.f[lev] = xcalloc(nl, sizeof(ocplx)); s.r[lev] = xcalloc(nl, sizeof(ocplx));
    }
    if (s.p.cycleopt == 1) v_cycle(&s, 0);
    if (s.p.cycleopt == 2) f_cycle(&s, 0);
    if (s.p.cycleopt == 3) w_cycle(&s, 0);
    for (lev = 1; lev < s.lmax; lev++) {
      emo_level_free(s.lev[lev]); free(s.u[lev]); free(s.f[lev]); free(s.r[lev]);
    }
  }
  memcpy(x, s.u[0], n * sizeof(ocplx));
  emo_curl_h(s.lev[0], s.omu, s.u[0], b);
  emo_level_free(s.lev[0]); free(s.u[0]); free(s.f[0]); free(s.r[0]);
  free(s.lev); free(s.u); free(s.f); free(s.r);
  return s.nhist;
}

/* ======================================================================= */
/* sources, receivers, gradient                                             */
/* ======================================================================= */

/* staggering per pattern: Ex Ey Ez use (s,n,n) (n,s,n) (n,n,s); the magnetic
 * patterns Hx Hy Hz use (n,s,s) (s,n,s) (s,s,n); s = staggered nodes */
static const int STAG[6][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}, {0, 1, 1}, {1, 0, 1}, {1, 1, 0}};

typedef struct { int i, j, k; float w1, w2, w3; const float *gx, *gy, *gz; } tri_t;

static tri_t locate(const emo_mg_t *g, int pat, float px, float py, float pz)
{
  tri_t t;
  t.gx = STAG[pat][0] ? g->x1s : g->x1;
  t.gy = STAG[pat][1] ? g->x2s : g->x2;
  t.gz = STAG[pat][2] ? g->x3s : g->x3;
  t.i = emo_find_index(g->n1 + 1, t.gx, px);
  t.j = emo_find_index(g->n2 + 1, t.gy, py);
  t.k = emo_find_index(g->n3 + 1, t.gz, pz);
  t.w1 = (px - t.gx[t.i]) / (t.gx[t.i + 1] - t.gx[t.i]);
  t.w2 = (py - t.gy[t.j]) / (t.gy[t.j + 1] - t.gy[t.j]);
  t.w3 = (pz - t.gz[t.k]) / (t.gz[t.k + 1] - t.gz[t.k]);
  return t;
}

static long vid(const emo_mg_t *g, int comp, int i, int j, int k)
{
  long n123 = (long)(g->n1 + 1) * (g->n2 + 1) * (g->n3 + 1);
  return i + (long)(g->n1 + 1) * (j + (long)(g->n2 + 1) * k) + comp * n123;
}

/* scatter `val` with trilinear weights onto the 8 corners (the order of the
 * eight updates of src/inject_extract.c:77-84) */
static void scatter8(const emo_mg_t *g, int comp, const tri_t *t, ocplx val, ocplx *v)
{
  v[vid(g, comp, t->i, t->j, t->k)] += val * (1. - t->w1) * (1. - t->w2) * (1. - t->w3);
  v[vid(g, comp, t->i + 1, t->j, t->k)] += val * t->w1 * (1. - t->w2) * (1. - t->w3);
  v[vid(g, comp, t->i, t->j + 1, t->k)] += val * (1. - t->w1) * t->w2 * (1. - t->w3);
  v[vid(g, comp, t->i + 1, t->j + 1, t->k)] += val * t->w1 * t->w2 * (1. - t->w3);
  v[vid(g, comp, t->i, t->j, t->k + 1)] += val * (1. - t->w1) * (1. - t->w2) * t->w3;
  v[vid(g, comp, t->i + 1, t->j, t->k + 1)] += val * t->w1 * (1. - t->w2) * t->w3;
  v[vid(g, comp, t->i, t->j + 1, t->k + 1)] += val * (1. - t->w1) * t->w2 * t->w3;
  v[vid(g, comp, t->i + 1, t->j + 1, t->k + 1)] += val * t->w1 * t->w2 * t->w3;
}

static ocplx gather8(const emo_mg_t *g, int comp, const tri_t *t, const ocplx *v)
{
  ocplx s = 0;
  s += v[vid(g, comp, t->i, t->j, t->k)] * (1. - t->w1) * (1. - t->w2) * (1. - t->w3);
  s += v[vid(g, comp, t->i + 1, t->j, t->k)] * t->w1 * (1. - t->w2) * (1. - t->w3);
  s += v[vid(g, comp, t->i, t->j + 1, t->k)] * (1. - t->w1) * t->w2 * (1. - t->w3);
  s += v[vid(g, comp, t->i + 1, t->j + 1, t->k)] * t->w1 * t->w2 * (1. - t->w3);
  s += v[vid(g, comp, t->i, t->j, t->k + 1)] * (1. - t->w1) * (1. - t->w2) * t->w3;
  s += v[vid(g, comp, t->i + 1, t->j, t->k + 1)] * t->w1 * (1. - t->w2) * t->w3;
  s += v[vid(g, comp, t->i, t->j + 1, t->k + 1)] * (1. - t->w1) * t->w2 * t->w3;
  s += v[vid(g, comp, t->i + 1, t->j + 1, t->k + 1)] * t->w1 * t->w2 * t->w3;
  return s;
}

static float tri_vol(const tri_t *t)
{
  return (t->gx[t->i + 1] - t->gx[t->i]) * (t->gy[t->j + 1] - t->gy[t->j]) * (t->gz[t->k + 1] - t->gz[t->k]);
}

/* dipole moment of the active channels rotated by azimuth/dip
 * (src/inject_extract.c:42-55), for electric (first = 0) or magnetic (3) */
static void moment(int nch, const int *ch, int first, ocplx amp_of_ch_dummy, const ocplx *amp,
                   float phi, float theta, ocplx *F)
{
  int ic;
  (void)amp_of_ch_dummy;
  F[0] = 0; F[1] = 0; F[2] = 0;
  for (ic = 0; ic < nch; ic++) {
    if (ch[ic] == first + 0) {
      F[0] += amp[ic] * cos(phi);
      F[1] += amp[ic] * sin(phi);
    } else if (ch[ic] == first + 1) {
      F[0] -= amp[ic] * sin(phi) * cos(theta);
      F[1] += amp[ic] * cos(phi) * cos(theta);
      F[2] += amp[ic] * sin(theta);
    } else if (ch[ic] == first + 2) {
      F[0] += amp[ic] * sin(phi) * sin(theta);
      F[1] -= amp[ic] * cos(phi) * sin(theta);
      F[2] += amp[ic] * cos(theta);
    }
  }
}

/* src/inject_extract.c:26-120 */
void emo_inject_source(const emo_mg_t *g, int nsrc, const float *sx, const float *sy, const float *sz,
                       const float *az, const float *dip, int nch, const int *ch, float omega, ocplx *b)
{
  long n = 3L * (g->n1 + 1) * (g->n2 + 1) * (g->n3 + 1);
  int is, comp;
  ocplx *amp = xcalloc(nch ? nch : 1, sizeof(ocplx));
  for (is = 0; is < nch; is++) amp[is] = 1;
  memset(b, 0, n * sizeof(ocplx));
  for (is = 0; is < nsrc; is++) {
    ocplx F[3];
    moment(nch, ch, 0, 0, amp, az[is], dip[is], F);
    for (comp = 0; comp < 3; comp++) {
      tri_t t = locate(g, comp, sx[is], sy[is], sz[is]);
      float vol = tri_vol(&t);
      ocplx sv = _Complex_I * omega * EMO_MU0 * F[comp] / vol;
      scatter8(g, comp, &t, sv, b);
    }
  }
  free(amp);
}

/* src/inject_extract.c:124-314 */
void emo_inject_adj_source(const emo_mg_t *g, int nrec, const float *rx, const float *ry, const float *rz,
                           const float *az, const float *dip, int nch, const int *ch,
                           const float _Complex *dres, float omega, ocplx *x, ocplx *b)
{
  long n = 3L * (g->n1 + 1) * (g->n2 + 1) * (g->n3 + 1);
  int ir, comp, ic, i, j, k;
  ocplx *amp = xcalloc(nch ? nch : 1, sizeof(ocplx));
  memset(b, 0, n * sizeof(ocplx));
  memset(x, 0, n * sizeof(ocplx));
  for (ir = 0; ir < nrec; ir++) {
    ocplx F[3];
    for (ic = 0; ic < nch; ic++) amp[ic] = dres[(long)ic * nrec + ir];
    moment(nch, ch, 0, 0, amp, az[ir], dip[ir], F);
    for (comp = 0; comp < 3; comp++) {
      tri_t t = locate(g, comp, rx[ir], ry[ir], rz[ir]);
      float vol = tri_vol(&t);
      ocplx rv = _Complex_I * omega * EMO_MU0 * F[comp] / vol;
      scatter8(g, comp, &t, rv, b);
    }
    moment(nch, ch, 3, 0, amp, az[ir], dip[ir], F);
    for (comp = 0; comp < 3; comp++) {
      tri_t t = locate(g, 3 + comp, rx[ir], ry[ir], rz[ir]);
      float vol = tri_vol(&t);
      ocplx rv = F[comp] / vol;
      scatter8(g, comp, &t, rv, x);
    }
  }
  /* b += curl M with the dual widths (mu_r = 1) */
  for (k = 1; k < g->n3 + 1; k++)
    for (j = 1; j < g->n2 + 1; j++)
      for (i = 1; i < g->n1 + 1; i++) {
        b[vid(g, 0, i, j, k)] += (x[vid(g, 2, i, j, k)] - x[vid(g, 2, i, j - 1, k)]) / g->d2[j] - (x[vid(g, 1, i, j, k)] - x[vid(g, 1, i, j, k - 1)]) / g->d3[k];
        b[vid(g, 1, i, j, k)] += (x[vid(g, 0, i, j, k)] - x[vid(g, 0, i, j, k - 1)]) / g->d3[k] - (x[vid(g, 2, i, j, k)] - x[vid(g, 2, i - 1, j, k)]) / g->d1[i];
        b[vid(g, 2, i, j, k)] += (x[vid(g, 1, i, j, k)] - x[vid(g, 1, i - 1, j, k)]) / g->d1[i] - (x[vid(g, 0, i, j, k)] - x[vid(g, 0, i, j - 1, k)]) / g->d2[j];
      }
  free(amp);
}

/* rotation into the receiver frame (src/inject_extract.c:391-399) */
static float _Complex rotate(int which, float phi, float theta, ocplx Fx, ocplx Fy, ocplx Fz)
{
  if (which == 0) return cos(phi) * Fx + sin(phi) * Fy;
  if (which == 1) return -sin(phi) * cos(theta) * Fx + cos(phi) * cos(theta) * Fy - sin(theta) * Fz;
  return sin(phi) * sin(theta) * Fx - cos(phi) * sin(theta) * Fy + cos(theta) * Fz;
}

/* src/inject_extract.c:317-474 */
void emo_extract_data(const emo_mg_t *g, int nrec, const float *rx, const float *ry, const float *rz,
                      const float *az, const float *dip, int nch, const int *ch,
                      const ocplx *x, const ocplx *b, float _Complex *dcal)
{
  int ir, ic;
  for (ir = 0; ir < nrec; ir++) {
    float phi = az[ir], theta = dip[ir];
    tri_t t;
    ocplx Fx, Fy, Fz;
    t = locate(g, 0, rx[ir], ry[ir], rz[ir]);
    Fx = gather8(g, 0, &t, x);
    t = locate(g, 1, rx[ir], ry[ir], rz[ir]);
    Fy = gather8(g, 1, &t, x);
    /* vertical current sigma33*Ez is continuous across layers; the cell
     * conductivity is taken unaveraged and converted with rho_water */
    t = locate(g, 2, rx[ir], ry[ir], rz[ir]);
    {
      const float *s = g->sigma33;
      int n1 = g->n1, n2 = g->n2;
#define S33(ii, jj, kk) s[(ii) + (long)n1 * ((jj) + (long)n2 * (kk))]
      Fz = 0;
      Fz += S33(t.i, t.j, t.k) * x[vid(g, 2, t.i, t.j, t.k)] * (1. - t.w1) * (1. - t.w2) * (1. - t.w3);
      Fz += S33(t.i + 1, t.j, t.k) * x[vid(g, 2, t.i + 1, t.j, t.k)] * t.w1 * (1. - t.w2) * (1. - t.w3);
      Fz += S33(t.i, t.j + 1, t.k) * x[vid(g, 2, t.i, t.j + 1, t.k)] * (1. - t.w1) * t.w2 * (1. - t.w3);
      Fz += S33(t.i + 1, t.j + 1, t.k) * x[vid(g, 2, t.i + 1, t.j + 1, t.k)] * t.w1 * t.w2 * (1. - t.w3);
      Fz += S33(t.i, t.j, t.k + 1) * x[vid(g, 2, t.i, t.j, t.k + 1)] * (1. - t.w1) * (1. - t.w2) * t.w3;
      Fz += S33(t.i + 1, t.j, t.k + 1) * x[vid(g, 2, t.i + 1, t.j, t.k + 1)] * t.w1 * (1. - t.w2) * t.w3;
      Fz += S33(t.i, t.j + 1, t.k + 1) * x[vid(g, 2, t.i, t.j + 1, t.k + 1)] * (1. - t.w1) * t.w2 * t.w3;
      Fz += S33(t.i + 1, t.j + 1, t.k + 1) * x[vid(g, 2, t.i + 1, t.j + 1, t.k + 1)] * t.w1 * t.w2 * t.w3;
#undef S33
      Fz *= g->rho_water;
    }
    for (ic = 0; ic < nch; ic++)
      if (ch[ic] <= 2) dcal[(long)ic * nrec + ir] = rotate(ch[ic], phi, theta, Fx, Fy, Fz);
    t = locate(g, 3, rx[ir], ry[ir], rz[ir]);
    Fx = gather8(g, 0, &t, b);
    t = locate(g, 4, rx[ir], ry[ir], rz[ir]);
    Fy = gather8(g, 1, &t, b);
    t = locate(g, 5, rx[ir], ry[ir], rz[ir]);
    Fz = gather8(g, 2, &t, b);
    for (ic = 0; ic < nch; ic++)
      if (ch[ic] >= 3) dcal[(long)ic * nrec + ir] = rotate(ch[ic] - 3, phi, theta, Fx, Fy, Fz);
  }
}

/* per-axis cell-centre interpolation tables (src/inject_extract.c:504-537) */
static void centre_tables(int n, float o, float d, int ng, const float *xg, const float *xgs,
                          int *in, float *wn, int *is, float *ws)
{
  int i, ii;
  for (i = 0; i < n; i++) {
    float c = o + (i + 0.5) * d;
    ii = emo_find_index(ng + 1, xg, c);
    wn[i] = (c - xg[ii]) / (xg[ii + 1] - xg[ii]);
    in[i] = ii;
    ii = emo_find_index(ng + 1, xgs, c);
    ws[i] = (c - xgs[ii]) / (xgs[ii + 1] - xgs[ii]);
    is[i] = ii;
  }
}

typedef struct { int *i1, *i2, *i3, *i1s, *i2s, *i3s; float *w1, *w2, *w3, *w1s, *w2s, *w3s; } ctab_t;

static void ctab_make(ctab_t *c, const emo_mg_t *g, int nx, int ny, int nz, float ox, float oy, float oz,
                      float dx, float dy, float dz)
{
  c->i1 = xcalloc(nx, sizeof(int)); c->i1s = xcalloc(nx, sizeof(int));
  c->i2 = xcalloc(ny, sizeof(int)); c->i2s = xcalloc(ny, sizeof(int));
  c->i3 = xcalloc(nz, sizeof(int)); c->i3s = xcalloc(nz, sizeof(int));
  c->w1 = xcalloc(nx, sizeof(float)); c->w1s = xcalloc(nx, sizeof(float));
  c->w2 = xcalloc(ny, sizeof(float)); c->w2s = xcalloc(ny, sizeof(float));
  c->w3 = xcalloc(nz, sizeof(float)); c->w3s = xcalloc(nz, sizeof(float));
  centre_tables(nx, ox, dx, g->n1, g->x1, g->x1s, c->i1, c->w1, c->i1s, c->w1s);
  centre_tables(ny, oy, dy, g->n2, g->x2, g->x2s, c->i2, c->w2, c->i2s, c->w2s);
  centre_tables(nz, oz, dz, g->n3, g->x3, g->x3s, c->i3, c->w3, c->i3s, c->w3s);
}

static void ctab_free(ctab_t *c)
{
  free(c->i1); free(c->i2); free(c->i3); free(c->i1s); free(c->i2s); free(c->i3s);
  free(c->w1); free(c->w2); free(c->w3); free(c->w1s); free(c->w2s); free(c->w3s);
}

static ocplx gather_w(const emo_mg_t *g, int comp, int ii, int jj, int kk, float wx, float wy, float wz, const ocplx *v)
{
  tri_t t;
  t.i = ii; t.j = jj; t.k = kk; t.w1 = wx; t.w2 = wy; t.w3 = wz;
  return gather8(g, comp, &t, v);
}

/* src/inject_extract.c:477-635 */
void emo_extract_efield(const emo_mg_t *g, int nx, int ny, int nz, float ox, float oy, float oz,
                        float dx, float dy, float dz, const float *rho33, const ocplx *x,
                        float _Complex *E1, float _Complex *E2, float _Complex *E3)
{
  ctab_t c;
  int i, j, k;
  ctab_make(&c, g, nx, ny, nz, ox, oy, oz, dx, dy, dz);
  for (k = 0; k < nz; k++)
    for (j = 0; j < ny; j++)
      for (i = 0; i < nx; i++) {
        long o = i + (long)nx * (j + (long)ny * k);
        E1[o] = gather_w(g, 0, c.i1s[i], c.i2[j], c.i3[k], c.w1s[i], c.w2[j], c.w3[k], x);
        E2[o] = gather_w(g, 1, c.i1[i], c.i2s[j], c.i3[k], c.w1[i], c.w2s[j], c.w3[k], x);
      }
  {
    const float *s = g->sigma33;
    int n1 = g->n1, n2 = g->n2;
#define S33(ii, jj, kk) s[(ii) + (long)n1 * ((jj) + (long)n2 * (kk))]
    for (k = 0; k < nz; k++) {
      int kk = c.i3s[k];
      float wz = c.w3s[k];
      for (j = 0; j < ny; j++) {
        int jj = c.i2[j];
        float wy = c.w2[j];
        for (i = 0; i < nx; i++) {
          int ii = c.i1[i];
          float wx = c.w1[i], sg;
          ocplx acc = 0;
          /* vertical current averaged over the four cells around each Ez edge */
          sg = 0.25 * (S33(ii, jj, kk) + S33(ii, jj - 1, kk) + S33(ii - 1, jj, kk) + S33(ii - 1, jj - 1, kk));
          acc += sg * x[vid(g, 2, ii, jj, kk)] * (1. - wx) * (1. - wy) * (1. - wz);
          sg = 0.25 * (S33(ii + 1, jj, kk) + S33(ii + 1, jj - 1, kk) + S33(ii, jj, kk) + S33(ii, jj - 1, kk));
          acc += sg * x[vid(g, 2, ii + 1, jj, kk)] * wx * (1. - wy) * (1. - wz);
          sg = 0.25 * (S33(ii, jj + 1, kk) + S33(ii, jj, kk) + S33(ii - 1, jj + 1, kk) + S33(ii - 1, jj, kk));
          acc += sg * x[vid(g, 2, ii, jj + 1, kk)] * (1. - wx) * wy * (1. - wz);
          sg = 0.25 * (S33(ii + 1, jj + 1, kk) + S33(ii + 1, jj, kk) + S33(ii, jj + 1, kk) + S33(ii, jj, kk));
          acc += sg * x[vid(g, 2, ii + 1, jj + 1, kk)] * wx * wy * (1. - wz);
          sg = 0.25 * (S33(ii, jj, kk + 1) + S33(ii, jj - 1, kk + 1) + S33(ii - 1, jj, kk + 1) + S33(ii - 1, jj - 1, kk + 1));
          acc += sg * x[vid(g, 2, ii, jj, kk + 1)] * (1. - wx) * (1. - wy) * wz;
          sg = 0.25 * (S33(ii + 1, jj, kk + 1) + S33(ii + 1, jj - 1, kk + 1) + S33(ii, jj, kk + 1) + S33(ii, jj - 1, kk + 1));
          acc += sg * x[vid(g, 2, ii + 1, jj, kk + 1)] * wx * (1. - wy) * wz;
          sg = 0.25 * (S33(ii, jj + 1, kk + 1) + S33(ii, jj, kk + 1) + S33(ii - 1, jj + 1, kk + 1) + S33(ii - 1, jj, kk + 1));
          acc += sg * x[vid(g, 2, ii, jj + 1, kk + 1)] * (1. - wx) * wy * wz;
          sg = 0.25 * (S33(ii + 1, jj + 1, kk + 1) + S33(ii + 1, jj, kk + 1) + S33(ii, jj + 1, kk + 1) + S33(ii, jj, kk + 1));
          acc += sg * x[vid(g, 2, ii + 1, jj + 1, kk + 1)] * wx * wy * wz;
          E3[i + (long)nx * (j + (long)ny * k)] = acc * rho33[i + (long)nx * (j + (long)ny * k)];
        }
      }
    }
#undef S33
  }
  ctab_free(&c);
}

/* src/inject_extract.c:638-785 */
void emo_extract_hfield(const emo_mg_t *g, int nx, int ny, int nz, float ox, float oy, float oz,
                        float dx, float dy, float dz, const ocplx *b,
                        float _Complex *H1, float _Complex *H2, float _Complex *H3)
{
  ctab_t c;
  int i, j, k;
  ctab_make(&c, g, nx, ny, nz, ox, oy, oz, dx, dy, dz);
  for (k = 0; k < nz; k++)
    for (j = 0; j < ny; j++)
      for (i = 0; i < nx; i++) {
        long o = i + (long)nx * (j + (long)ny * k);
        H1[o] = gather_w(g, 0, c.i1[i], c.i2s[j], c.i3s[k], c.w1[i], c.w2s[j], c.w3s[k], b);
        H2[o] = gather_w(g, 1, c.i1s[i], c.i2[j], c.i3s[k], c.w1s[i], c.w2[j], c.w3s[k], b);
        H3[o] = gather_w(g, 2, c.i1s[i], c.i2s[j], c.i3[k], c.w1s[i], c.w2s[j], c.w3[k], b);
      }
  ctab_free(&c);
}

/* mute box around a point: cell index range clipped to the model
 * (src/build_gradient.c:64-111) */
static void mute_box(const emo_gradspec_t *s, float px, float py, float pz, float r, float *grad)
{
  int i1min = (px - r - s->ox) / s->dx, i1max = (px + r - s->ox) / s->dx;
  int i2min = (py - r - s->oy) / s->dy, i2max = (py + r - s->oy) / s->dy;
  int i3min = (pz - r - s->oz) / s->dz, i3max = (pz + r - s->oz) / s->dz;
  int k, i1, i2, i3;
#define CLIP(v, n) ((v) < 0 ? 0 : ((v) > (n) - 1 ? (n) - 1 : (v)))
  i1min = CLIP(i1min, s->nx); i1max = CLIP(i1max, s->nx);
  i2min = CLIP(i2min, s->ny); i2max = CLIP(i2max, s->ny);
  i3min = CLIP(i3min, s->nz); i3max = CLIP(i3max, s->nz);
#undef CLIP
  for (k = 0; k < s->npar; k++)
    for (i3 = i3min; i3 <= i3max; i3++)
      for (i2 = i2min; i2 <= i2max; i2++)
        for (i1 = i1min; i1 <= i1max; i1++)
          grad[i1 + (long)s->nx * (i2 + (long)s->ny * (i3 + (long)s->nz * k))] = 0.;
}

/* src/build_gradient.c:13-111 */
void emo_build_gradient(const emo_gradspec_t *s, int nfreq,
                        const float _Complex *E1f, const float _Complex *E2f, const float _Complex *E3f,
                        const float _Complex *E1a, const float _Complex *E2a, const float _Complex *E3a,
                        int nsrc, const float *sx, const float *sy, const float *sz,
                        int nrec, const float *rx, const float *ry, const float *rz,
                        float *grad, float *phess)
{
  long nxyz = (long)s->nx * s->ny * s->nz;
  float vol = s->dx * s->dy * s->dz;
  int k, i1, i2, i3, f, is;
  for (k = 0; k < s->npar; k++)
    for (i3 = 0; i3 < s->nz; i3++)
      for (i2 = 0; i2 < s->ny; i2++)
        for (i1 = 0; i1 < s->nx; i1++) {
          long c = i1 + (long)s->nx * (i2 + (long)s->ny * i3), o = c + nxyz * k;
          grad[o] = 0;
          for (f = 0; f < nfreq; f++) {
            long fc = c + nxyz * f;
            if (s->idxpar[k] == 1) {
              grad[o] += creal(E1f[fc] * E1a[fc]);
              grad[o] += creal(E2f[fc] * E2a[fc]);
            }
            if (s->idxpar[k] == 2) grad[o] += creal(E3f[fc] * E3a[fc]);
          }
          if (s->idxpar[k] == 1) grad[o] *= vol / s->rho11[c];
          if (s->idxpar[k] == 2) grad[o] *= vol / s->rho33[c];
          if (s->oz + (i3 + 0.5) * s->dz <= s->bathy[i1 + (long)s->nx * i2] + s->dz) grad[o] = 0.;
          if (s->preco == 1 && phess) {
            phess[o] = 0.;
            for (f = 0; f < nfreq; f++) {
              long fc = c + nxyz * f;
              if (s->idxpar[k] == 1) {
                phess[o] += cabs(E1f[fc]) * cabs(E1f[fc]);
                phess[o] += cabs(E2f[fc]) * cabs(E2f[fc]);
              }
              if (s->idxpar[k] == 2) phess[o] += cabs(E3f[fc]) * cabs(E3f[fc]);
            }
            if (s->idxpar[k] == 1) phess[o] *= vol / s->rho11[c];
            if (s->idxpar[k] == 2) phess[o] *= vol / s->rho33[c];
          }
        }
  for (is = 0; is < nsrc; is++) mute_box(s, sx[is], sy[is], sz[is], s->dsmute, grad);
  for (is = 0; is < nrec; is++) mute_box(s, rx[is], ry[is], rz[is], s->drmute, grad);
}
