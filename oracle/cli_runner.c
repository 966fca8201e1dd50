/* cli_runner.c -- test-only driver over the reference's host sources
 * (src/cstd.c, src/emf_init_free.c, src/acq_init_free.c,
 * src/read_write_data.c, src/cal_uncertainty_noise.c, src/frannor.c,
 * src/regularization.c, src/triangle_smoothing.c, src/precondition.c,
 * src/fg_mod_reg.c, compiled in place by oracle/Makefile target `cli`)
 * that prints what they compute, for the golden fixtures of the
 * command-line program.
 *
 * This file is test infrastructure: never linked into the product, its
 * binary lives in oracle/_ref/ (git-ignored).
 *
 * Tasks (key=value, getpar conventions):
 *   task=rannor n=N [seed=S]     N deviates of frannor (after srannor(S))
 *   task=pars keys=k1,k2 types=i,f  getpar of each key, as the reference
 *                                reads it
 *   task=init rank=R nranks=P     emf_init + acq_init of rank R of P
 *   task=reg|smooth|precond|modreg n1= n2= n3= fx= ...  regularisation,
 *                                triangle smoothing, preconditioning and
 *                                fg_mod_reg on volumes read from files
 *   task=noise rank=R nranks=P    emf_init, acq_init, read_data,
 *                                cal_uncertainty_noise, write_data(obs),
 *                                then the misfit of src/fg_fwi.c:207-225
 *                                against the noise-free data, write_rmse
 * Floats are printed as C99 hex floats so they compare bit for bit.
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <math.h>
#include <complex.h>
#include "cstd.h"
#include "emf.h"
#include "acq.h"
#include "fwi.h"

int iproc = 0, nproc = 1, ierr = 0;

void emf_init(emf_t *emf);
void acq_init(acq_t *acq, emf_t *emf);
void read_data(acq_t *acq, emf_t *emf);
void write_data(acq_t *acq, emf_t *emf, char *fname, _Complex float ***dcal_fd);
void write_rmse(acq_t *acq, emf_t *emf);
float cal_uncertainty_noise(acq_t *acq, emf_t *emf);
float frannor(void);
void srannor(int seed);
float regularization_tikhonov(float *x, float *g, int n1, int n2, int n3, float d1, float d2, float d3);
float regularization_tv(float *x, float *g, int n1, int n2, int n3, float d1, float d2, float d3);
void triangle_smoothing(float ***mod, int n1, int n2, int n3, int r1, int r2, int r3, int repeat);
void precondition(emf_t *emf, fwi_t *fwi, float *g);
void fg_mod_reg(acq_t *acq, emf_t *emf, fwi_t *fwi, float *x, float *g);

static void farr(const char *name, const float *a, int n);

/* n floats from the file named by parameter key */
static float *read_floats(const char *key, size_t n)
{
  char *name;
  FILE *fp;
  float *a = alloc1float(n);
  if (!getparstring(key, &name)) err("need %s=", key);
  fp = fopen(name, "rb");
  if (!fp || fread(a, sizeof(float), n, fp) != n) err("cannot read %zu floats from %s", n, name);
  fclose(fp);
  return a;
}

/* the model-side functions of the inversion on volumes read from files:
 * regularisation (task=reg), smoothing (task=smooth), preconditioning
 * (task=precond preco=1|2|3) and fg_mod_reg (task=modreg) */
static void task_model_side(const char *task)
{
  int n1, n2, n3, npar;
  size_t nxyz, n;
  float *x, *g, f;
  if (!getparint("n1", &n1) || !getparint("n2", &n2) || !getparint("n3", &n3)) err("need n1 n2 n3");
  if (!getparint("npar", &npar)) npar = 1;
  nxyz = (size_t)n1*n2*n3;
  n = nxyz*npar;
  x = read_floats("fx", n);
  if (strcmp(task, "reg") == 0) {
    g = alloc1float(nxyz);
    f = regularization_tikhonov(x, g, n1, n2, n3, 1, 1, 1);
    printf("tik %a\n", f); farr("gtik", g, nxyz);
    f = regularization_tv(x, g, n1, n2, n3, 1, 1, 1);
    printf("tv %a\n", f); farr("gtv", g, nxyz);
  } else if (strcmp(task, "smooth") == 0) {
    int r1, r2, r3, rep;
    float ***m = alloc3float(n1, n2, n3);
    if (!getparint("r1", &r1)) r1 = 3;
    if (!getparint("r2", &r2)) r2 = 3;
    if (!getparint("r3", &r3)) r3 = 1;
    if (!getparint("repeat", &rep)) rep = 1;
    memcpy(&m[0][0][0], x, nxyz*sizeof(float));
    triangle_smoothing(m, n1, n2, n3, r1, r2, r3, rep);
    farr("smooth", &m[0][0][0], nxyz);
  } else {
    emf_t e;
    fwi_t w;
    acq_t a;
    float freq;
    memset(&e, 0, sizeof e); memset(&w, 0, sizeof w); memset(&a, 0, sizeof a);
    e.nx = n1; e.ny = n2; e.nz = n3;
    if (!getparfloat("dx", &e.dx)) e.dx = 100;
    if (!getparfloat("dy", &e.dy)) e.dy = 100;
    if (!getparfloat("dz", &e.dz)) e.dz = 50;
    if (!getparfloat("oz", &e.oz)) e.oz = 0;
    if (!getparfloat("freq", &freq)) freq = 0.5;
    e.omegas = alloc1float(1);
    e.omegas[0] = 2.*PI*freq;
    e.bathy = alloc2float(n1, n2);
    memcpy(&e.bathy[0][0], read_floats("fbathy", (size_t)n1*n2), (size_t)n1*n2*sizeof(float));
    w.npar = npar; w.n = n;
    if (!getparint("preco", &w.preco)) w.preco = 1;
    if (!getparint("r1", &w.r1)) w.r1 = 3;
    if (!getparint("r2", &w.r2)) w.r2 = 3;
    if (!getparint("r3", &w.r3)) w.r3 = 1;
    if (!getparint("repeat", &w.repeat)) w.repeat = 1;
    if (!getparint("iter", &w.iter)) w.iter = 0;
    if (!getparfloat("gamma1", &w.gamma1)) w.gamma1 = 10;
    if (!getparfloat("gamma2", &w.gamma2)) w.gamma2 = 0;
    w.grad = alloc4float(n1, n2, n3, npar);
    w.phess = alloc4float(n1, n2, n3, npar);
    memcpy(&w.phess[0][0][0][0], read_floats("fphess", n), n*sizeof(float));
    w.xref = read_floats("fxref", n);
    g = read_floats("fg", n);
    if (strcmp(task, "precond") == 0) {
      precondition(&e, &w, g);
    } else {
      fg_mod_reg(&a, &e, &w, x, g);
      printf("fcost_mod %a\n", w.fcost_mod);
    }
    farr("g", g, n);
  }
}

static void task_rannor(void)
{
  int n, seed, i;
  if (!getparint("n", &n)) n = 1000;
  if (getparint("seed", &seed)) srannor(seed);
  for (i = 0; i < n; i++) printf("%a\n", frannor());
}

static void task_pars(void)
{
  int nk = countparval("keys"), k, j, r;
  char **keys, **types;
  if (nk == 0 || countparval("types") != nk) err("keys= and types= of equal length");
  keys = (char **)alloc1(nk, sizeof(char *));
  types = (char **)alloc1(nk, sizeof(char *));
  getparstringarray("keys", keys);
  getparstringarray("types", types);
  for (k = 0; k < nk; k++) {
    printf("%s %s", keys[k], types[k]);
    if (types[k][0] == 'i') {
      int v[64];
      r = getparint(keys[k], v);
      printf(" %d %d", r, countparval(keys[k]));
      for (j = 0; j < r; j++) printf(" %d", v[j]);
    } else if (types[k][0] == 'f') {
      float v[64];
      r = getparfloat(keys[k], v);
      printf(" %d %d", r, countparval(keys[k]));
      for (j = 0; j < r; j++) printf(" %a", v[j]);
    } else if (types[k][0] == 'd') {
      double v[64];
      r = getpardouble(keys[k], v);
      printf(" %d %d", r, countparval(keys[k]));
      for (j = 0; j < r; j++) printf(" %a", v[j]);
    } else if (types[k][0] == 's') {
      char *v = NULL;
      r = getparstring(keys[k], &v);
      printf(" %d %d [%s]", r, countparval(keys[k]), r ? v : "");
    } else {
      char *v[64];
      r = getparstringarray(keys[k], v);
      printf(" %d %d", r, countparval(keys[k]));
      for (j = 0; j < r; j++) printf(" [%s]", v[j]);
    }
    printf("\n");
  }
}

static void farr(const char *name, const float *a, int n)
{
  int i;
  printf("%s %d", name, n);
  for (i = 0; i < n; i++) printf(" %a", a[i]);
  printf("\n");
}

static void dump_init(acq_t *acq, emf_t *emf)
{
  int i;
  printf("nfreq %d\n", emf->nfreq);
  farr("freqs", emf->freqs, emf->nfreq);
  farr("omegas", emf->omegas, emf->nfreq);
  printf("chsrc %d", emf->nchsrc);
  for (i = 0; i < emf->nchsrc; i++) printf(" %s", emf->chsrc[i]);
  printf("\nchrec %d", emf->nchrec);
  for (i = 0; i < emf->nchrec; i++) printf(" %s", emf->chrec[i]);
  printf("\nints %d %d %d %d %d %d %d %d\n", emf->reciprocity, emf->nx, emf->ny, emf->nz, emf->n1, emf->n2,
         emf->n3, emf->nb);
  printf("floats %a %a %a %a %a %a %a %a %a %a %a %a %a\n", emf->dx, emf->dy, emf->dz, emf->ox, emf->oy,
         emf->oz, emf->d1min, emf->d2min, emf->d3min, emf->lextend, emf->rho_air, emf->rho_water, 0.0f);
  printf("acq_ints %d %d %d %d %d %d %d\n", acq->nsrc, acq->nrec, acq->nsubsrc, acq->nsubrec, acq->nsrc_total,
         acq->nrec_total, acq->shot_idx[iproc]);
  printf("acq_floats %a %a %a %a %a %a %a %a\n", acq->lensrc, acq->lenrec, acq->x1min, acq->x1max, acq->x2min,
         acq->x2max, acq->x3min, acq->x3max);
  farr("src_x1", acq->src_x1, acq->nsrc);
  farr("src_x2", acq->src_x2, acq->nsrc);
  farr("src_x3", acq->src_x3, acq->nsrc);
  farr("src_azimuth", acq->src_azimuth, acq->nsrc);
  farr("src_dip", acq->src_dip, acq->nsrc);
  farr("rec_x1", acq->rec_x1, acq->nrec);
  farr("rec_x2", acq->rec_x2, acq->nrec);
  farr("rec_x3", acq->rec_x3, acq->nrec);
  farr("rec_azimuth", acq->rec_azimuth, acq->nrec);
  farr("rec_dip", acq->rec_dip, acq->nrec);
}

/* src/fg_fwi.c:55-64 (the noise parameters) and :207-225 (the misfit) */
static void task_noise(acq_t *acq, emf_t *emf)
{
  int ic, ifr, ir, nd = emf->nchrec * emf->nfreq * acq->nrec;
  float mse, fcost = 0, s, s1, s2;
  char fname[64];
  _Complex float ***clean;
  if (!getparint("addnoise", &emf->addnoise)) emf->addnoise = 1;
  if (!getparfloat("amp_perc", &emf->amp_perc)) emf->amp_perc = 0.03;
  if (!getparfloat("delta_phi", &emf->delta_phi)) emf->delta_phi = 0;
  if (!getparfloat("delta_theta", &emf->delta_theta)) emf->delta_theta = 0;
  if (!getparfloat("noisefloorE", &emf->noisefloorE)) emf->noisefloorE = 1e-15;
  if (!getparfloat("noisefloorH", &emf->noisefloorH)) emf->noisefloorH = 1e-13;
  emf->dobs_fd = alloc3complexf(acq->nrec, emf->nfreq, emf->nchrec);
  emf->dcal_fd = alloc3complexf(acq->nrec, emf->nfreq, emf->nchrec);
  emf->dres_fd = alloc3complexf(acq->nrec, emf->nfreq, emf->nchrec);
  emf->delta_emf = alloc3float(acq->nrec, emf->nfreq, emf->nchrec);
  emf->rmse = alloc3float(acq->nrec, emf->nfreq, emf->nchrec);
  clean = alloc3complexf(acq->nrec, emf->nfreq, emf->nchrec);
  memset(&emf->dobs_fd[0][0][0], 0, nd * sizeof(_Complex float));
  memset(&emf->dres_fd[0][0][0], 0, nd * sizeof(_Complex float));
  read_data(acq, emf);
  memcpy(&clean[0][0][0], &emf->dobs_fd[0][0][0], nd * sizeof(_Complex float));
  mse = cal_uncertainty_noise(acq, emf);
  sprintf(fname, "obs_%04d.txt", acq->shot_idx[iproc]);
  write_data(acq, emf, fname, emf->dobs_fd);
  printf("mse %a\n", mse);
  farr("delta", &emf->delta_emf[0][0][0], nd);
  memcpy(&emf->dcal_fd[0][0][0], &clean[0][0][0], nd * sizeof(_Complex float));
  memset(&emf->rmse[0][0][0], 0, nd * sizeof(float));
  for (ic = 0; ic < emf->nchrec; ic++)
    for (ifr = 0; ifr < emf->nfreq; ifr++) {
      s = 5.5 * sqrt(2. * emf->rho_water / (emf->omegas[ifr] * mu0));
      for (ir = 0; ir < acq->nrec; ir++) {
        s1 = acq->rec_x1[ir] - acq->src_x1[0];
        s2 = acq->rec_x2[ir] - acq->src_x2[0];
        if (sqrt(s1 * s1 + s2 * s2) > s) {
          emf->dres_fd[ic][ifr][ir] = emf->dobs_fd[ic][ifr][ir] - emf->dcal_fd[ic][ifr][ir];
          emf->dres_fd[ic][ifr][ir] /= emf->delta_emf[ic][ifr][ir];
          emf->rmse[ic][ifr][ir] = cabs(emf->dres_fd[ic][ifr][ir]);
          fcost += emf->dres_fd[ic][ifr][ir] * conj(emf->dres_fd[ic][ifr][ir]);
          emf->dres_fd[ic][ifr][ir] /= emf->delta_emf[ic][ifr][ir];
          emf->dres_fd[ic][ifr][ir] = conj(emf->dres_fd[ic][ifr][ir]);
        }
      }
    }
  fcost *= 0.5;
  write_rmse(acq, emf);
  printf("fcost %a\n", fcost);
  farr("dres", (float *)&emf->dres_fd[0][0][0], 2 * nd);
}

int main(int argc, char **argv)
{
  char *task;
  emf_t emf;
  acq_t acq;
  initargs(argc, argv);
  if (!getparstring("task", &task)) task = "init";
  if (strcmp(task, "rannor") == 0) { task_rannor(); return 0; }
  if (strcmp(task, "pars") == 0) { task_pars(); return 0; }
  if (!strcmp(task, "reg") || !strcmp(task, "smooth") || !strcmp(task, "precond") || !strcmp(task, "modreg")) {
    task_model_side(task);
    return 0;
  }
  if (!getparint("nranks", &nproc)) nproc = 1;
  if (!getparint("rank", &iproc)) iproc = 0;
  memset(&emf, 0, sizeof emf);
  memset(&acq, 0, sizeof acq);
  if (!getparint("verb", &emf.verb)) emf.verb = 0;
  emf_init(&emf);
  acq_init(&acq, &emf);
  printf("---- dump\n");
  if (strcmp(task, "noise") == 0) task_noise(&acq, &emf);
  else dump_init(&acq, &emf);
  return 0;
}
