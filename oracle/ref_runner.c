/* ref_runner.c -- test-only driver that links the reference's own hot-path
 * sources (compiled in place from the reference tree by oracle/Makefile) and
 * dumps their outputs as raw little-endian binaries for the golden fixtures.
 *
 * This file is test infrastructure.  It is never linked into the product
 * library and never shipped as part of it; its output binary lives in
 * oracle/_ref/ (git-ignored).
 *
 * It replaces the reference entry point src/main.c:24-84 (which needs MPI)
 * for a single rank: iproc=0, nproc=1.  The modelling task follows the call
 * sequence of src/do_modelling.c:28-104, the gradient task the one of
 * src/fg_fwi.c:139-270 without the MPI sums (one shot per process).
 *
 * Built with -DEMMI_DROPIN (oracle/Makefile target `dropin`), the same
 * driver links the reference's host sources (getpar, emf_init, acq_init,
 * write_data) against libemmi_mgfd.so in place of the reference's hot-path
 * files: the link-time swap a libEMMI_MGFD host program makes to use the
 * GPU library.  The unit task needs the reference's own src/gmg.c and is
 * left out of that build.
 *
 * Tasks (key=value on the command line, getpar conventions):
 *   task=model  outdir=DIR [dumpfields=1]   mode-0 pipeline
 *   task=units  outdir=DIR [seed=N]         single-kernel known answers
 *   task=grad   outdir=DIR                  mode-2 gradient pipeline
 *   task=fwi    (in the run directory)      one rank of a mode-2 run, with
 *                                            the reference's file outputs
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <math.h>
#include <complex.h>
#include <sys/stat.h>
#include <time.h>
#include "cstd.h"
#include "emf.h"
#include "acq.h"
#include "fwi.h"

/* globals the reference's main.c normally defines (one rank) */
int iproc = 0, nproc = 1, ierr = 0;

/* reference entry points (declared locally, as its own callers do) */
void emf_init(emf_t *emf);
void emf_free(emf_t *emf);
void acq_init(acq_t *acq, emf_t *emf);
void acq_free(acq_t *acq);
void regrid_init(acq_t *acq, emf_t *emf);
void regrid_free(emf_t *emf);
void gmg_init(emf_t *emf, int ifreq);
void gmg_free(void);
void gmg_apply(int n, double _Complex *b, double _Complex *x);
void inject_source(acq_t *acq, emf_t *emf, double _Complex *b, int ifreq);
void inject_adj_source(acq_t *acq, emf_t *emf, double _Complex *x, double _Complex *b, int ifreq);
void extract_emf_data(acq_t *acq, emf_t *emf, double _Complex *x, double _Complex *b, int ifreq);
void extract_electric_field(acq_t *acq, emf_t *emf, double _Complex *x, int ifreq);
void extract_magnetic_field(acq_t *acq, emf_t *emf, double _Complex *b, int ifreq);
void write_data(acq_t *acq, emf_t *emf, char *fname, _Complex float ***dcal_fd);
void read_data(acq_t *acq, emf_t *emf);
void write_rmse(acq_t *acq, emf_t *emf);
float cal_uncertainty_noise(acq_t *acq, emf_t *emf);
void build_gradient(acq_t *acq, emf_t *emf, fwi_t *fwi,
                    _Complex float ****E1f, _Complex float ****E2f, _Complex float ****E3f,
                    _Complex float ****E1a, _Complex float ****E2a, _Complex float ****E3a);
int find_index(int n, float *x, float val);

#ifndef EMMI_DROPIN
/* Layout of one multigrid level as src/gmg.c:24-34 declares it; needed only
 * by the unit task to fill and read the reference's level arrays. */
typedef struct {
  int n1, n2, n3;
  double *x1, *x2, *x3;
  double *x1s, *x2s, *x3s;
  double *d1, *d2, *d3;
  double *d1s, *d2s, *d3s;
  double _Complex ****u, ****f, ****r;
  double ***sigma11, ***sigma22, ***sigma33;
  double ***invmur;
  int sc[3];
} ref_level_t;
extern ref_level_t *gmg;   /* src/gmg.c:35 */
extern int icycle;         /* src/gmg.c:15 */
void grid_init(ref_level_t *g, int lev);
void grid_free(ref_level_t *g, int lev);
void residual(ref_level_t *g, int lev);
void restriction(ref_level_t *g, int lev);
void prolongation(ref_level_t *g, int lev);
void line_gauss_seidel_x(ref_level_t *g, int lev, int iter);
void line_gauss_seidel_y(ref_level_t *g, int lev, int iter);
void line_gauss_seidel_z(ref_level_t *g, int lev, int iter);
void gauss_seidel(ref_level_t *g, int lev, int iter);
void compute_H_from_E(ref_level_t *g);
#endif

static char *outdir;

static void dump(const char *name, const void *p, size_t bytes)
{
  char path[4096];
  FILE *fp;
  snprintf(path, sizeof(path), "%s/%s.bin", outdir, name);
  fp = fopen(path, "wb");
  if (!fp) { fprintf(stderr, "cannot write %s\n", path); exit(1); }
  if (bytes && fwrite(p, 1, bytes, fp) != bytes) { fprintf(stderr, "short write %s\n", path); exit(1); }
  fclose(fp);
}

static void dump_grid(emf_t *emf)
{
  int nx = emf->nx + 2*emf->nb, ny = emf->ny + 2*emf->nb, nz = emf->nz + 2*emf->nb;
  size_t n1 = emf->n1 + 1, n2 = emf->n2 + 1, n3 = emf->n3 + 1;
  size_t nc = (size_t)emf->n1*emf->n2*emf->n3;
  dump("xx", emf->xx, (nx+1)*sizeof(float));
  dump("yy", emf->yy, (ny+1)*sizeof(float));
  dump("zz", emf->zz, (nz+1)*sizeof(float));
  dump("x1", emf->x1, n1*sizeof(float));  dump("x2", emf->x2, n2*sizeof(float));  dump("x3", emf->x3, n3*sizeof(float));
  dump("x1s", emf->x1s, n1*sizeof(float)); dump("x2s", emf->x2s, n2*sizeof(float)); dump("x3s", emf->x3s, n3*sizeof(float));
  dump("d1", emf->d1, n1*sizeof(float));  dump("d2", emf->d2, n2*sizeof(float));  dump("d3", emf->d3, n3*sizeof(float));
  dump("d1s", emf->d1s, n1*sizeof(float)); dump("d2s", emf->d2s, n2*sizeof(float)); dump("d3s", emf->d3s, n3*sizeof(float));
  dump("sigma11", &emf->sigma11[0][0][0], nc*sizeof(float));
  dump("sigma22", &emf->sigma22[0][0][0], nc*sizeof(float));
  dump("sigma33", &emf->sigma33[0][0][0], nc*sizeof(float));
}

/* cell indices and trilinear weights the injector/extractor use, per receiver
 * and staggering pattern: for each receiver 6 patterns x (i,j,k) ints and
 * 6 x (w1,w2,w3) floats, the patterns ordered Ex, Ey, Ez, Hx, Hy, Hz. */
static void dump_rec_index(acq_t *acq, emf_t *emf)
{
  int nrec = acq->nrec, irec, p;
  int *idx = malloc(sizeof(int)*18*(nrec+1));
  float *w = malloc(sizeof(float)*18*(nrec+1));
  for (irec = 0; irec <= nrec; irec++) {
    float px, py, pz;
    if (irec < nrec) { px = acq->rec_x1[irec]; py = acq->rec_x2[irec]; pz = acq->rec_x3[irec]; }
    else { px = acq->src_x1[0]; py = acq->src_x2[0]; pz = acq->src_x3[0]; }
    for (p = 0; p < 6; p++) {
      /* staggering per pattern: (x staggered?, y?, z?) */
      static const int st[6][3] = {{1,0,0},{0,1,0},{0,0,1},{0,1,1},{1,0,1},{1,1,0}};
      float *gx = st[p][0] ? emf->x1s : emf->x1;
      float *gy = st[p][1] ? emf->x2s : emf->x2;
      float *gz = st[p][2] ? emf->x3s : emf->x3;
      int i = find_index(emf->n1+1, gx, px);
      int j = find_index(emf->n2+1, gy, py);
      int k = find_index(emf->n3+1, gz, pz);
      idx[18*irec + 3*p + 0] = i; idx[18*irec + 3*p + 1] = j; idx[18*irec + 3*p + 2] = k;
      w[18*irec + 3*p + 0] = (px - gx[i])/(gx[i+1] - gx[i]);
      w[18*irec + 3*p + 1] = (py - gy[j])/(gy[j+1] - gy[j]);
      w[18*irec + 3*p + 2] = (pz - gz[k])/(gz[k+1] - gz[k]);
    }
  }
  dump("rec_idx", idx, sizeof(int)*18*(nrec+1));
  dump("rec_w", w, sizeof(float)*18*(nrec+1));
  free(idx); free(w);
}

static unsigned long long rng_state = 88172645463325252ULL;
static double urand(void)
{
  rng_state ^= rng_state << 13; rng_state ^= rng_state >> 7; rng_state ^= rng_state << 17;
  return (double)(rng_state >> 11) * (1.0/9007199254740992.0) - 0.5;
}

#ifndef EMMI_DROPIN
/* fill a level vector with random values on interior edges, zero elsewhere */
static void fill_interior(ref_level_t *g, double _Complex ****v)
{
  int i, j, k;
  for (k = 0; k <= g->n3; k++)
    for (j = 0; j <= g->n2; j++)
      for (i = 0; i <= g->n1; i++) {
        v[0][k][j][i] = (i < g->n1 && j > 0 && j < g->n2 && k > 0 && k < g->n3) ? urand() + I*urand() : 0;
        v[1][k][j][i] = (j < g->n2 && i > 0 && i < g->n1 && k > 0 && k < g->n3) ? urand() + I*urand() : 0;
        v[2][k][j][i] = (k < g->n3 && i > 0 && i < g->n1 && j > 0 && j < g->n2) ? urand() + I*urand() : 0;
      }
}

static size_t level_len(ref_level_t *g) { return 3*(size_t)(g->n1+1)*(g->n2+1)*(g->n3+1); }

static void task_units(acq_t *acq, emf_t *emf)
{
  int seed, c, it;
  size_t n;
  double _Complex *u0;
  char name[64];
  if (!getparint("seed", &seed)) seed = 1;
  rng_state += 0x9E3779B97F4A7C15ULL*(unsigned long long)seed;

  regrid_init(acq, emf);
  dump_grid(emf);
  gmg_init(emf, 0);
  icycle = 0;
  grid_init(gmg, 0);
  n = level_len(&gmg[0]);
  dump("lev0_sigma11", &gmg[0].sigma11[0][0][0], sizeof(double)*gmg[0].n1*gmg[0].n2*gmg[0].n3);
  dump("lev0_sigma22", &gmg[0].sigma22[0][0][0], sizeof(double)*gmg[0].n1*gmg[0].n2*gmg[0].n3);
  dump("lev0_sigma33", &gmg[0].sigma33[0][0][0], sizeof(double)*gmg[0].n1*gmg[0].n2*gmg[0].n3);

  /* residual r = f - A u */
  fill_interior(&gmg[0], gmg[0].u);
  fill_interior(&gmg[0], gmg[0].f);
  u0 = malloc(n*sizeof(double _Complex));
  memcpy(u0, &gmg[0].u[0][0][0][0], n*sizeof(double _Complex));
  dump("u", u0, n*sizeof(double _Complex));
  dump("f", &gmg[0].f[0][0][0][0], n*sizeof(double _Complex));
  residual(gmg, 0);
  dump("r", &gmg[0].r[0][0][0][0], n*sizeof(double _Complex));
  compute_H_from_E(gmg);
  dump("H", &gmg[0].f[0][0][0][0], n*sizeof(double _Complex));

  /* smoothers from the same (u, f) */
  fill_interior(&gmg[0], gmg[0].f);
  dump("f2", &gmg[0].f[0][0][0][0], n*sizeof(double _Complex));
  for (it = 0; it < 2; it++) {
    memcpy(&gmg[0].u[0][0][0][0], u0, n*sizeof(double _Complex));
    line_gauss_seidel_x(gmg, 0, it);
    snprintf(name, sizeof name, "lgs_x_it%d", it); dump(name, &gmg[0].u[0][0][0][0], n*sizeof(double _Complex));
    memcpy(&gmg[0].u[0][0][0][0], u0, n*sizeof(double _Complex));
    line_gauss_seidel_y(gmg, 0, it);
    snprintf(name, sizeof name, "lgs_y_it%d", it); dump(name, &gmg[0].u[0][0][0][0], n*sizeof(double _Complex));
    memcpy(&gmg[0].u[0][0][0][0], u0, n*sizeof(double _Complex));
    line_gauss_seidel_z(gmg, 0, it);
    snprintf(name, sizeof name, "lgs_z_it%d", it); dump(name, &gmg[0].u[0][0][0][0], n*sizeof(double _Complex));
    memcpy(&gmg[0].u[0][0][0][0], u0, n*sizeof(double _Complex));
    gauss_seidel(gmg, 0, it);
    snprintf(name, sizeof name, "pgs_it%d", it); dump(name, &gmg[0].u[0][0][0][0], n*sizeof(double _Complex));
  }

  /* inter-grid transfer for each semi-coarsening pattern */
  for (c = 0; c < 3; c++) {
    size_t nc1;
    icycle = c;
    grid_init(gmg, 1);
    nc1 = level_len(&gmg[1]);
    fill_interior(&gmg[0], gmg[0].r);
    snprintf(name, sizeof name, "rf_c%d", c); dump(name, &gmg[0].r[0][0][0][0], n*sizeof(double _Complex));
    restriction(gmg, 0);
    snprintf(name, sizeof name, "fc_c%d", c); dump(name, &gmg[1].f[0][0][0][0], nc1*sizeof(double _Complex));
    fill_interior(&gmg[1], gmg[1].u);
    snprintf(name, sizeof name, "uc_c%d", c); dump(name, &gmg[1].u[0][0][0][0], nc1*sizeof(double _Complex));
    memcpy(&gmg[0].u[0][0][0][0], u0, n*sizeof(double _Complex));
    prolongation(gmg, 0);
    snprintf(name, sizeof name, "uf_c%d", c); dump(name, &gmg[0].u[0][0][0][0], n*sizeof(double _Complex));
    snprintf(name, sizeof name, "lev1_sigma11_c%d", c);
    dump(name, &gmg[1].sigma11[0][0][0], sizeof(double)*gmg[1].n1*gmg[1].n2*gmg[1].n3);
    snprintf(name, sizeof name, "lev1_sigma33_c%d", c);
    dump(name, &gmg[1].sigma33[0][0][0], sizeof(double)*gmg[1].n1*gmg[1].n2*gmg[1].n3);
    {
      int dims[3] = {gmg[1].n1, gmg[1].n2, gmg[1].n3};
      snprintf(name, sizeof name, "lev1_dims_c%d", c); dump(name, dims, sizeof dims);
    }
    grid_free(gmg, 1);
  }
  free(u0);
  grid_free(gmg, 0);
  gmg_free();
  regrid_free(emf);
}

#endif

static void task_model(acq_t *acq, emf_t *emf)
{
  int ifreq, n, icheck, dumpfields;
  double _Complex *x, *b;
  char name[64];
  if (!getparint("icheck", &icheck)) icheck = 0;
  if (!getparint("dumpfields", &dumpfields)) dumpfields = 0;
  if (icheck) {
    emf->E1 = alloc4complexf(emf->nx, emf->ny, emf->nz, emf->nfreq);
    emf->E2 = alloc4complexf(emf->nx, emf->ny, emf->nz, emf->nfreq);
    emf->E3 = alloc4complexf(emf->nx, emf->ny, emf->nz, emf->nfreq);
    emf->H1 = alloc4complexf(emf->nx, emf->ny, emf->nz, emf->nfreq);
    emf->H2 = alloc4complexf(emf->nx, emf->ny, emf->nz, emf->nfreq);
    emf->H3 = alloc4complexf(emf->nx, emf->ny, emf->nz, emf->nfreq);
  }
  emf->dcal_fd = alloc3complexf(acq->nrec, emf->nfreq, emf->nchrec);
  n = 3*(emf->n1+1)*(emf->n2+1)*(emf->n3+1);
  x = alloc1complex(n);
  b = alloc1complex(n);
  memset(b, 0, n*sizeof(double _Complex));
  regrid_init(acq, emf);
  dump_grid(emf);
  dump_rec_index(acq, emf);
  for (ifreq = 0; ifreq < emf->nfreq; ifreq++) {
    gmg_init(emf, ifreq);
    inject_source(acq, emf, b, ifreq);
    if (dumpfields) { snprintf(name, sizeof name, "b_f%d", ifreq); dump(name, b, n*sizeof(double _Complex)); }
    memset(x, 0, n*sizeof(double _Complex));
    {
      struct timespec t0, t1;
      printf("solve_begin ifreq=%d\n", ifreq);
      clock_gettime(CLOCK_MONOTONIC, &t0);
      gmg_apply(n, b, x);
      clock_gettime(CLOCK_MONOTONIC, &t1);
      printf("solve_end ifreq=%d seconds=%.6f\n", ifreq, (t1.tv_sec - t0.tv_sec) + 1e-9 * (t1.tv_nsec - t0.tv_nsec));
    }
    if (dumpfields) {
      snprintf(name, sizeof name, "E_f%d", ifreq); dump(name, x, n*sizeof(double _Complex));
      snprintf(name, sizeof name, "H_f%d", ifreq); dump(name, b, n*sizeof(double _Complex));
    }
    extract_emf_data(acq, emf, x, b, ifreq);
    if (icheck) {
      extract_electric_field(acq, emf, x, ifreq);
      extract_magnetic_field(acq, emf, b, ifreq);
    }
    gmg_free();
  }
  free1complex(x);
  free1complex(b);
  regrid_free(emf);
  dump("dcal", &emf->dcal_fd[0][0][0], sizeof(_Complex float)*acq->nrec*emf->nfreq*emf->nchrec);
  {
    char path[4096];
    snprintf(path, sizeof path, "%s/emf_%04d.txt", outdir, acq->shot_idx[iproc]);
    write_data(acq, emf, path, emf->dcal_fd);
  }
  if (icheck) {
    size_t nv = (size_t)emf->nfreq*emf->nx*emf->ny*emf->nz*sizeof(_Complex float);
    dump("E1", &emf->E1[0][0][0][0], nv); dump("E2", &emf->E2[0][0][0][0], nv); dump("E3", &emf->E3[0][0][0][0], nv);
    dump("H1", &emf->H1[0][0][0][0], nv); dump("H2", &emf->H2[0][0][0][0], nv); dump("H3", &emf->H3[0][0][0][0], nv);
  }
}

/* Gradient of one shot: src/fg_fwi.c:139-270 with the observed data read from
 * dobs= (complex64 [nchrec][nfreq][nrec]) and unit weights delta= (float32,
 * same shape), no noise, no MPI sum, no regularisation. */
static void task_grad(acq_t *acq, emf_t *emf)
{
  int ifreq, n, ichrec, irec, npar = 2, k;
  double _Complex *x, *b;
  _Complex float ****E1f, ****E2f, ****E3f, ****E1a, ****E2a, ****E3a;
  fwi_t fwi;
  char *fdobs, *fdelta, *fbathy;
  FILE *fp;
  float fcost = 0, s, s1, s2;
  size_t nd = (size_t)emf->nchrec*emf->nfreq*acq->nrec;

  memset(&fwi, 0, sizeof fwi);
  if (!getparstring("dobs", &fdobs)) err("need dobs=");
  if (!getparstring("delta", &fdelta)) err("need delta=");
  if (!getparstring("fbathy", &fbathy)) err("need fbathy=");
  if (!getparfloat("dsmute", &emf->dsmute)) emf->dsmute = 800;
  if (!getparfloat("drmute", &emf->drmute)) emf->drmute = 200;
  if (!getparint("preco", &fwi.preco)) fwi.preco = 1;
  fwi.npar = npar;
  fwi.idxpar = alloc1int(npar); fwi.idxpar[0] = 1; fwi.idxpar[1] = 2;
  fwi.n = npar*emf->nx*emf->ny*emf->nz;
  fwi.grad = alloc4float(emf->nx, emf->ny, emf->nz, npar);
  fwi.phess = alloc4float(emf->nx, emf->ny, emf->nz, npar);
  emf->bathy = alloc2float(emf->nx, emf->ny);
  fp = fopen(fbathy, "rb"); if (!fp) err("bathy");
  if (fread(&emf->bathy[0][0], sizeof(float), emf->nx*emf->ny, fp) != (size_t)(emf->nx*emf->ny)) err("bathy size");
  fclose(fp);
  emf->dobs_fd = alloc3complexf(acq->nrec, emf->nfreq, emf->nchrec);
  emf->dcal_fd = alloc3complexf(acq->nrec, emf->nfreq, emf->nchrec);
  emf->dres_fd = alloc3complexf(acq->nrec, emf->nfreq, emf->nchrec);
  emf->delta_emf = alloc3float(acq->nrec, emf->nfreq, emf->nchrec);
  memset(&emf->dres_fd[0][0][0], 0, nd*sizeof(_Complex float));
  fp = fopen(fdobs, "rb"); if (!fp) err("dobs");
  if (fread(&emf->dobs_fd[0][0][0], sizeof(_Complex float), nd, fp) != nd) err("dobs size");
  fclose(fp);
  fp = fopen(fdelta, "rb"); if (!fp) err("delta");
  if (fread(&emf->delta_emf[0][0][0], sizeof(float), nd, fp) != nd) err("delta size");
  fclose(fp);

  E1f = alloc4complexf(emf->nx, emf->ny, emf->nz, emf->nfreq);
  E2f = alloc4complexf(emf->nx, emf->ny, emf->nz, emf->nfreq);
  E3f = alloc4complexf(emf->nx, emf->ny, emf->nz, emf->nfreq);
  E1a = alloc4complexf(emf->nx, emf->ny, emf->nz, emf->nfreq);
  E2a = alloc4complexf(emf->nx, emf->ny, emf->nz, emf->nfreq);
  E3a = alloc4complexf(emf->nx, emf->ny, emf->nz, emf->nfreq);

  n = 3*(emf->n1+1)*(emf->n2+1)*(emf->n3+1);
  emf->E1 = E1f; emf->E2 = E2f; emf->E3 = E3f;
  regrid_init(acq, emf);
  for (ifreq = 0; ifreq < emf->nfreq; ifreq++) {
    gmg_init(emf, ifreq);
    x = alloc1complex(n); b = alloc1complex(n);
    inject_source(acq, emf, b, ifreq);
    memset(x, 0, n*sizeof(double _Complex));
    gmg_apply(n, b, x);
    extract_emf_data(acq, emf, x, b, ifreq);
    extract_electric_field(acq, emf, x, ifreq);
    free1complex(x); free1complex(b);
    gmg_free();
  }
  regrid_free(emf);
  dump("dcal", &emf->dcal_fd[0][0][0], nd*sizeof(_Complex float));

  for (ichrec = 0; ichrec < emf->nchrec; ichrec++)
    for (ifreq = 0; ifreq < emf->nfreq; ifreq++) {
      s = 5.5*sqrt(2.*emf->rho_water/(emf->omegas[ifreq]*mu0));
      for (irec = 0; irec < acq->nrec; irec++) {
        s1 = acq->rec_x1[irec] - acq->src_x1[0];
        s2 = acq->rec_x2[irec] - acq->src_x2[0];
        if (sqrt(s1*s1 + s2*s2) > s) {
          emf->dres_fd[ichrec][ifreq][irec] = emf->dobs_fd[ichrec][ifreq][irec] - emf->dcal_fd[ichrec][ifreq][irec];
          emf->dres_fd[ichrec][ifreq][irec] /= emf->delta_emf[ichrec][ifreq][irec];
          fcost += emf->dres_fd[ichrec][ifreq][irec]*conj(emf->dres_fd[ichrec][ifreq][irec]);
          emf->dres_fd[ichrec][ifreq][irec] /= emf->delta_emf[ichrec][ifreq][irec];
          emf->dres_fd[ichrec][ifreq][irec] = conj(emf->dres_fd[ichrec][ifreq][irec]);
        }
      }
    }
  fcost *= 0.5;
  dump("fcost", &fcost, sizeof fcost);
  dump("dres", &emf->dres_fd[0][0][0], nd*sizeof(_Complex float));

  emf->E1 = E1a; emf->E2 = E2a; emf->E3 = E3a;
  regrid_init(acq, emf);
  for (ifreq = 0; ifreq < emf->nfreq; ifreq++) {
    gmg_init(emf, ifreq);
    x = alloc1complex(n); b = alloc1complex(n);
    inject_adj_source(acq, emf, x, b, ifreq);
    if (ifreq == 0) dump("badj_f0", b, n*sizeof(double _Complex));
    memset(x, 0, n*sizeof(double _Complex));
    gmg_apply(n, b, x);
    extract_emf_data(acq, emf, x, b, ifreq);
    extract_electric_field(acq, emf, x, ifreq);
    free1complex(x); free1complex(b);
    gmg_free();
  }
  regrid_free(emf);
  build_gradient(acq, emf, &fwi, E1f, E2f, E3f, E1a, E2a, E3a);
  {
    size_t nv = (size_t)emf->nfreq*emf->nx*emf->ny*emf->nz*sizeof(_Complex float);
    dump("E1f", &E1f[0][0][0][0], nv); dump("E2f", &E2f[0][0][0][0], nv); dump("E3f", &E3f[0][0][0][0], nv);
    dump("E1a", &E1a[0][0][0][0], nv); dump("E2a", &E2a[0][0][0][0], nv); dump("E3a", &E3a[0][0][0][0], nv);
  }
  for (k = 0; k < npar; k++) (void)k;
  dump("grad", &fwi.grad[0][0][0][0], sizeof(float)*fwi.n);
  dump("phess", &fwi.phess[0][0][0][0], sizeof(float)*fwi.n);
}

/* One gradient evaluation of one rank as src/fg_fwi.c runs it: fg_fwi_init
 * (:43-137: bathymetry, read_data, cal_uncertainty_noise, obs_<shot>.txt)
 * and fg_fwi (:139-285 up to the data gradient: the log/exp round trip of
 * the model from src/do_fwi.c:110-121, forward solves, syn_<shot>.txt, the
 * misfit, rmse_<shot>.txt, adjoint solves, build_gradient, and the files
 * gradient_dat_Rv/_Rh), without MPI (one rank: the sums are copies) and
 * without the regularisation that follows the file writes.  The muted
 * residuals start at zero. */
static void task_fwi(acq_t *acq, emf_t *emf)
{
  int ifreq, n, ic, ir, j, i, npar;
  size_t nxyz = (size_t)emf->nx*emf->ny*emf->nz, nd;
  double _Complex *x, *b;
  _Complex float ****E1f, ****E2f, ****E3f, ****E1a, ****E2a, ****E3a;
  fwi_t fwi;
  char *fbathy, fname[64];
  FILE *fp;
  float tmp, fcost = 0, s, s1, s2, *xx;

  memset(&fwi, 0, sizeof fwi);
  if (!getparstring("fbathy", &fbathy)) err("Need fbathy= for CSEM inversion");
  if (!getparint("addnoise", &emf->addnoise)) emf->addnoise = 1;
  if (!getparfloat("amp_perc", &emf->amp_perc)) emf->amp_perc = 0.03;
  if (!getparfloat("delta_phi", &emf->delta_phi)) emf->delta_phi = 0;
  if (!getparfloat("delta_theta", &emf->delta_theta)) emf->delta_theta = 0;
  if (!getparfloat("noisefloorE", &emf->noisefloorE)) emf->noisefloorE = 1e-15;
  if (!getparfloat("noisefloorH", &emf->noisefloorH)) emf->noisefloorH = 1e-13;
  if (!getparfloat("dsmute", &emf->dsmute)) emf->dsmute = 800;
  if (!getparfloat("drmute", &emf->drmute)) emf->drmute = 200;
  if (!getparint("preco", &fwi.preco)) fwi.preco = 1;
  if (!getparint("npar", &npar)) npar = 2;
  fwi.npar = npar;
  fwi.idxpar = alloc1int(npar);
  for (j = 0; j < npar; j++) fwi.idxpar[j] = j + 1;
  getparint("idxpar", fwi.idxpar);
  fwi.n = npar*nxyz;

  emf->bathy = alloc2float(emf->nx, emf->ny);
  fp = fopen(fbathy, "rb"); if (!fp) err("bathy");
  if (fread(&emf->bathy[0][0], sizeof(float), emf->nx*emf->ny, fp) != (size_t)(emf->nx*emf->ny)) err("bathy size");
  fclose(fp);
  emf->depthmax = emf->bathy[0][0];
  for (i = 0; i < emf->nx*emf->ny; i++) emf->depthmax = MAX(emf->bathy[0][i], emf->depthmax);
  if (emf->depthmax < acq->src_x3[0]) err("coordinate error: source placed below seabed!");

  nd = (size_t)emf->nchrec*emf->nfreq*acq->nrec;
  emf->dobs_fd = alloc3complexf(acq->nrec, emf->nfreq, emf->nchrec);
  emf->dcal_fd = alloc3complexf(acq->nrec, emf->nfreq, emf->nchrec);
  emf->dres_fd = alloc3complexf(acq->nrec, emf->nfreq, emf->nchrec);
  emf->delta_emf = alloc3float(acq->nrec, emf->nfreq, emf->nchrec);
  emf->rmse = alloc3float(acq->nrec, emf->nfreq, emf->nchrec);
  memset(&emf->dobs_fd[0][0][0], 0, nd*sizeof(_Complex float));
  memset(&emf->dres_fd[0][0][0], 0, nd*sizeof(_Complex float));
  read_data(acq, emf);
  tmp = cal_uncertainty_noise(acq, emf);
  printf("addnoise=%d, target rmse=%e\n", emf->addnoise, sqrt(tmp/nd));
  sprintf(fname, "obs_%04d.txt", acq->shot_idx[iproc]);
  write_data(acq, emf, fname, emf->dobs_fd);

  fwi.grad = alloc4float(emf->nx, emf->ny, emf->nz, npar);
  fwi.phess = alloc4float(emf->nx, emf->ny, emf->nz, npar);
  xx = alloc1float(fwi.n);
  for (j = 0; j < npar; j++)
    for (i = 0; i < (int)nxyz; i++) {
      if (fwi.idxpar[j] == 1) xx[i + nxyz*j] = log(emf->rho11[0][0][i]);
      if (fwi.idxpar[j] == 2) xx[i + nxyz*j] = log(emf->rho33[0][0][i]);
    }
  for (j = 0; j < npar; j++)
    for (i = 0; i < (int)nxyz; i++) {
      s = exp(xx[i + nxyz*j]);
      if (fwi.idxpar[j] == 1) { emf->rho11[0][0][i] = s; emf->rho22[0][0][i] = s; }
      if (fwi.idxpar[j] == 2) emf->rho33[0][0][i] = s;
    }

  E1f = alloc4complexf(emf->nx, emf->ny, emf->nz, emf->nfreq);
  E2f = alloc4complexf(emf->nx, emf->ny, emf->nz, emf->nfreq);
  E3f = alloc4complexf(emf->nx, emf->ny, emf->nz, emf->nfreq);
  E1a = alloc4complexf(emf->nx, emf->ny, emf->nz, emf->nfreq);
  E2a = alloc4complexf(emf->nx, emf->ny, emf->nz, emf->nfreq);
  E3a = alloc4complexf(emf->nx, emf->ny, emf->nz, emf->nfreq);
  n = 3*(emf->n1+1)*(emf->n2+1)*(emf->n3+1);
  emf->E1 = E1f; emf->E2 = E2f; emf->E3 = E3f;
  regrid_init(acq, emf);
  for (ifreq = 0; ifreq < emf->nfreq; ifreq++) {
    gmg_init(emf, ifreq);
    x = alloc1complex(n); b = alloc1complex(n);
    inject_source(acq, emf, b, ifreq);
    memset(x, 0, n*sizeof(double _Complex));
    gmg_apply(n, b, x);
    extract_emf_data(acq, emf, x, b, ifreq);
    extract_electric_field(acq, emf, x, ifreq);
    free1complex(x); free1complex(b);
    gmg_free();
  }
  regrid_free(emf);
  sprintf(fname, "syn_%04d.txt", acq->shot_idx[iproc]);
  write_data(acq, emf, fname, emf->dcal_fd);

  memset(&emf->rmse[0][0][0], 0, nd*sizeof(float));
  for (ic = 0; ic < emf->nchrec; ic++)
    for (ifreq = 0; ifreq < emf->nfreq; ifreq++) {
      s = 5.5*sqrt(2.*emf->rho_water/(emf->omegas[ifreq]*mu0));
      for (ir = 0; ir < acq->nrec; ir++) {
        s1 = acq->rec_x1[ir] - acq->src_x1[0];
        s2 = acq->rec_x2[ir] - acq->src_x2[0];
        if (sqrt(s1*s1 + s2*s2) > s) {
          emf->dres_fd[ic][ifreq][ir] = emf->dobs_fd[ic][ifreq][ir] - emf->dcal_fd[ic][ifreq][ir];
          emf->dres_fd[ic][ifreq][ir] /= emf->delta_emf[ic][ifreq][ir];
          emf->rmse[ic][ifreq][ir] = cabs(emf->dres_fd[ic][ifreq][ir]);
          fcost += emf->dres_fd[ic][ifreq][ir]*conj(emf->dres_fd[ic][ifreq][ir]);
          emf->dres_fd[ic][ifreq][ir] /= emf->delta_emf[ic][ifreq][ir];
          emf->dres_fd[ic][ifreq][ir] = conj(emf->dres_fd[ic][ifreq][ir]);
        }
      }
    }
  fcost *= 0.5;
  write_rmse(acq, emf);
  printf("isrc=%d, fcost=%g, rmse=%g\n", acq->shot_idx[iproc], fcost, sqrt(2.*fcost/nd));
  dump("fcost", &fcost, sizeof fcost);

  emf->E1 = E1a; emf->E2 = E2a; emf->E3 = E3a;
  regrid_init(acq, emf);
  for (ifreq = 0; ifreq < emf->nfreq; ifreq++) {
    gmg_init(emf, ifreq);
    x = alloc1complex(n); b = alloc1complex(n);
    inject_adj_source(acq, emf, x, b, ifreq);
    memset(x, 0, n*sizeof(double _Complex));
    gmg_apply(n, b, x);
    extract_emf_data(acq, emf, x, b, ifreq);
    extract_electric_field(acq, emf, x, ifreq);
    free1complex(x); free1complex(b);
    gmg_free();
  }
  regrid_free(emf);
  build_gradient(acq, emf, &fwi, E1f, E2f, E3f, E1a, E2a, E3a);
  fp = fopen("gradient_dat_Rv", "wb");
  fwrite(&fwi.grad[0][0][0][0], nxyz*sizeof(float), 1, fp);
  fclose(fp);
  if (npar > 1) {
    fp = fopen("gradient_dat_Rh", "wb");
    fwrite(&fwi.grad[0][0][0][0] + nxyz, nxyz*sizeof(float), 1, fp);
    fclose(fp);
  }
}

int main(int argc, char **argv)
{
  emf_t *emf;
  acq_t *acq;
  char *task;
  setvbuf(stdout, NULL, _IOLBF, 0);
  initargs(argc, argv);
  if (!getparstring("outdir", &outdir)) outdir = ".";
  mkdir(outdir, 0755);
  if (!getparstring("task", &task)) task = "model";
  acq = calloc(1, sizeof(acq_t));
  emf = calloc(1, sizeof(emf_t));
  if (!getparint("mode", &emf->mode)) emf->mode = 0;
  if (!getparint("verb", &emf->verb)) emf->verb = 1;
  emf_init(emf);
  acq_init(acq, emf);
#ifdef EMMI_DROPIN
  if (strcmp(task, "units") == 0) { fprintf(stderr, "task=units needs the reference's src/gmg.c\n"); exit(1); }
#else
  if (strcmp(task, "units") == 0) task_units(acq, emf);
#endif
  else if (strcmp(task, "grad") == 0) task_grad(acq, emf);
  else if (strcmp(task, "fwi") == 0) task_fwi(acq, emf);
  else task_model(acq, emf);
  printf("ref_runner done\n");
  return 0;
}
