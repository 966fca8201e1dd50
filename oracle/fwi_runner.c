/* fwi_runner.c -- test-only driver: the reference's inversion on one rank.
 *
 * The reference's own src/do_fwi.c, src/opt.c, src/fg_mod_reg.c,
 * src/regularization.c, src/precondition.c, src/triangle_smoothing.c and its
 * host and hot-path sources are compiled in place (oracle/Makefile target
 * `fwi`).  src/fg_fwi.c includes <mpi.h>, which this image lacks, so its
 * four functions are restated below for one rank, where each MPI_Allreduce
 * (a sum over ranks) is a copy: fg_fwi_init (src/fg_fwi.c:43-120),
 * fg_fwi_free (:122-137), fg_fwi (:139-305), Hv_fwi (:307-311, empty in the
 * reference).  Muted residuals start at zero.
 *
 * This file is test infrastructure: never linked into the product, its
 * binary lives in oracle/_ref/ (git-ignored).  It takes the reference's
 * command line (mode=1 or mode=2, shots=k for the one rank).
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
void do_fwi(acq_t *acq, emf_t *emf);
void read_data(acq_t *acq, emf_t *emf);
void write_rmse(acq_t *acq, emf_t *emf);
void write_data(acq_t *acq, emf_t *emf, char *fname, _Complex float ***dcal_fd);
float cal_uncertainty_noise(acq_t *acq, emf_t *emf);
void regrid_init(acq_t *acq, emf_t *emf);
void regrid_free(emf_t *emf);
void gmg_init(emf_t *emf, int ifreq);
void gmg_free(void);
void gmg_apply(int n, double _Complex *b, double _Complex *x);
void inject_source(acq_t *acq, emf_t *emf, double _Complex *b, int ifreq);
void inject_adj_source(acq_t *acq, emf_t *emf, double _Complex *x, double _Complex *b, int ifreq);
void extract_emf_data(acq_t *acq, emf_t *emf, double _Complex *x, double _Complex *b, int ifreq);
void extract_electric_field(acq_t *acq, emf_t *emf, double _Complex *x, int ifreq);
void build_gradient(acq_t *acq, emf_t *emf, fwi_t *fwi,
                    _Complex float ****E1f, _Complex float ****E2f, _Complex float ****E3f,
                    _Complex float ****E1a, _Complex float ****E2a, _Complex float ****E3a);
void fg_mod_reg(acq_t *acq, emf_t *emf, fwi_t *fwi, float *x, float *g);

static acq_t *acq;
static emf_t *emf;
static fwi_t *fwi;

void fg_fwi_init(acq_t *acq_, emf_t *emf_, fwi_t *fwi_)
{
  char *fbathy, fname[64];
  FILE *fp;
  int i;
  size_t nd;
  acq = acq_; emf = emf_; fwi = fwi_;
  if (!getparstring("fbathy", &fbathy)) err("Need fbathy= for CSEM inversion");
  if (!getparint("addnoise", &emf->addnoise)) emf->addnoise = 1;
  if (!getparfloat("amp_perc", &emf->amp_perc)) emf->amp_perc = 0.03;
  if (!getparfloat("delta_phi", &emf->delta_phi)) emf->delta_phi = 0;
  if (!getparfloat("delta_theta", &emf->delta_theta)) emf->delta_theta = 0;
  if (!getparfloat("noisefloorE", &emf->noisefloorE)) emf->noisefloorE = 1e-15;
  if (!getparfloat("noisefloorH", &emf->noisefloorH)) emf->noisefloorH = 1e-13;
  if (!getparfloat("dsmute", &emf->dsmute)) emf->dsmute = 800;
  if (!getparfloat("drmute", &emf->drmute)) emf->drmute = 200;
  if (!getparfloat("gamma1", &fwi->gamma1)) fwi->gamma1 = 10.;
  if (!getparfloat("gamma2", &fwi->gamma2)) fwi->gamma2 = 0.;
  if (!getparint("r1", &fwi->r1)) fwi->r1 = 3;
  if (!getparint("r2", &fwi->r2)) fwi->r2 = 3;
  if (!getparint("r3", &fwi->r3)) fwi->r3 = 1;
  if (!getparint("repeat", &fwi->repeat)) fwi->repeat = 1;
  emf->bathy = alloc2float(emf->nx, emf->ny);
  nd = (size_t)emf->nchrec*emf->nfreq*acq->nrec;
  emf->dobs_fd = alloc3complexf(acq->nrec, emf->nfreq, emf->nchrec);
  emf->dcal_fd = alloc3complexf(acq->nrec, emf->nfreq, emf->nchrec);
  emf->dres_fd = alloc3complexf(acq->nrec, emf->nfreq, emf->nchrec);
  emf->delta_emf = alloc3float(acq->nrec, emf->nfreq, emf->nchrec);
  emf->rmse = alloc3float(acq->nrec, emf->nfreq, emf->nchrec);
  memset(&emf->dobs_fd[0][0][0], 0, nd*sizeof(_Complex float));
  memset(&emf->dres_fd[0][0][0], 0, nd*sizeof(_Complex float));
  fp = fopen(fbathy, "rb");
  if (fp == NULL) err("cannot open file fbathy=%s\n", fbathy);
  if (fread(&emf->bathy[0][0], sizeof(float), emf->nx*emf->ny, fp) != (size_t)(emf->nx*emf->ny))
    err("size parameter does not match the file-bathy!");
  fclose(fp);
  emf->depthmax = emf->bathy[0][0];
  for (i = 0; i < emf->nx*emf->ny; i++) emf->depthmax = MAX(emf->bathy[0][i], emf->depthmax);
  if (emf->depthmax < acq->src_x3[0]) err("coordinate error: source placed below seabed!");
  read_data(acq, emf);
  fwi->mse = cal_uncertainty_noise(acq, emf);   /* MPI_Allreduce over one rank */
  sprintf(fname, "obs_%04d.txt", acq->shot_idx[iproc]);
  write_data(acq, emf, fname, emf->dobs_fd);
  fwi->xref = alloc1float(fwi->n);
  fwi->grad = alloc4float(emf->nx, emf->ny, emf->nz, fwi->npar);
  if (fwi->preco == 1) fwi->phess = alloc4float(emf->nx, emf->ny, emf->nz, fwi->npar);
  fwi->iter = 0;
  fwi->firstgrad = 1;
  fwi->alpha = 1;
}

void fg_fwi_free()
{
  free2float(emf->bathy);
  free3complexf(emf->dobs_fd);
  free3complexf(emf->dcal_fd);
  free3complexf(emf->dres_fd);
  free3float(emf->delta_emf);
  free3float(emf->rmse);
  free1float(fwi->xref);
  free4float(fwi->grad);
  if (fwi->preco == 1) free4float(fwi->phess);
}

static void solve_all(_Complex float ****E1, _Complex float ****E2, _Complex float ****E3, int adjoint)
{
  int ifreq, n = 3*(emf->n1+1)*(emf->n2+1)*(emf->n3+1);
  double _Complex *x, *b;
  emf->E1 = E1; emf->E2 = E2; emf->E3 = E3;
  regrid_init(acq, emf);
  for (ifreq = 0; ifreq < emf->nfreq; ifreq++) {
    gmg_init(emf, ifreq);
    x = alloc1complex(n); b = alloc1complex(n);
    if (adjoint) inject_adj_source(acq, emf, x, b, ifreq);
    else inject_source(acq, emf, b, ifreq);
    memset(x, 0, n*sizeof(double _Complex));
    gmg_apply(n, b, x);
    extract_emf_data(acq, emf, x, b, ifreq);
    extract_electric_field(acq, emf, x, ifreq);
    free1complex(x); free1complex(b);
    gmg_free();
  }
  regrid_free(emf);
}

float fg_fwi(float *xx, float *g)
{
  float fcost, s, s1, s2;
  int i1, i2, i3, j, k, n, ifreq, ichrec, irec;
  _Complex float ****E1f, ****E2f, ****E3f, ****E1a, ****E2a, ****E3a;
  char fname[64];
  FILE *fp;

  emf->rhomin = exp(xx[0]);
  emf->rhomax = exp(xx[0]);
  for (j = 0; j < fwi->npar; j++)
    for (i3 = 0; i3 < emf->nz; i3++)
      for (i2 = 0; i2 < emf->ny; i2++)
        for (i1 = 0; i1 < emf->nx; i1++) {
          k = i1 + emf->nx*(i2 + emf->ny*(i3 + emf->nz*j));
          s = exp(xx[k]);
          if (fwi->idxpar[j] == 1) { emf->rho11[i3][i2][i1] = s; emf->rho22[i3][i2][i1] = s; }
          if (fwi->idxpar[j] == 2) emf->rho33[i3][i2][i1] = s;
          emf->rhomin = MIN(emf->rhomin, s);
          emf->rhomax = MAX(emf->rhomax, s);
        }
  E1f = alloc4complexf(emf->nx, emf->ny, emf->nz, emf->nfreq);
  E2f = alloc4complexf(emf->nx, emf->ny, emf->nz, emf->nfreq);
  E3f = alloc4complexf(emf->nx, emf->ny, emf->nz, emf->nfreq);
  E1a = alloc4complexf(emf->nx, emf->ny, emf->nz, emf->nfreq);
  E2a = alloc4complexf(emf->nx, emf->ny, emf->nz, emf->nfreq);
  E3a = alloc4complexf(emf->nx, emf->ny, emf->nz, emf->nfreq);
  solve_all(E1f, E2f, E3f, 0);
  sprintf(fname, "syn_%04d.txt", acq->shot_idx[iproc]);
  write_data(acq, emf, fname, emf->dcal_fd);

  memset(&emf->rmse[0][0][0], 0, acq->nrec*emf->nfreq*emf->nchrec*sizeof(float));
  fcost = 0.;
  for (ichrec = 0; ichrec < emf->nchrec; ichrec++)
    for (ifreq = 0; ifreq < emf->nfreq; ifreq++) {
      s = 5.5*sqrt(2.*emf->rho_water/(emf->omegas[ifreq]*mu0));
      for (irec = 0; irec < acq->nrec; irec++) {
        s1 = acq->rec_x1[irec] - acq->src_x1[0];
        s2 = acq->rec_x2[irec] - acq->src_x2[0];
        if (sqrt(s1*s1 + s2*s2) > s) {
          emf->dres_fd[ichrec][ifreq][irec] = emf->dobs_fd[ichrec][ifreq][irec] - emf->dcal_fd[ichrec][ifreq][irec];
          emf->dres_fd[ichrec][ifreq][irec] /= emf->delta_emf[ichrec][ifreq][irec];
          emf->rmse[ichrec][ifreq][irec] = cabs(emf->dres_fd[ichrec][ifreq][irec]);
          fcost += emf->dres_fd[ichrec][ifreq][irec]*conj(emf->dres_fd[ichrec][ifreq][irec]);
          emf->dres_fd[ichrec][ifreq][irec] /= emf->delta_emf[ichrec][ifreq][irec];
          emf->dres_fd[ichrec][ifreq][irec] = conj(emf->dres_fd[ichrec][ifreq][irec]);
        }
      }
    }
  fcost *= 0.5;
  write_rmse(acq, emf);
  n = emf->nchrec*emf->nfreq*acq->nrec;
  printf("isrc=%d, fcost=%g, rmse=%g\n", acq->shot_idx[iproc], fcost, sqrt(2.*fcost/n));
  fwi->fcost_dat = fcost;                        /* MPI_Allreduce over one rank */
  n *= nproc;
  fwi->mse = sqrt(2.*fwi->fcost_dat/n);

  solve_all(E1a, E2a, E3a, 1);
  build_gradient(acq, emf, fwi, E1f, E2f, E3f, E1a, E2a, E3a);
  memcpy(g, &fwi->grad[0][0][0][0], fwi->n*sizeof(float));   /* MPI_Allreduce over one rank */
  free4complexf(E1f); free4complexf(E2f); free4complexf(E3f);
  free4complexf(E1a); free4complexf(E2a); free4complexf(E3a);

  if (emf->verb) {
    n = emf->nx*emf->ny*emf->nz;
    fp = fopen("gradient_dat_Rv", "wb"); fwrite(g, n*sizeof(float), 1, fp); fclose(fp);
    fp = fopen("gradient_dat_Rh", "wb"); fwrite(g + n, n*sizeof(float), 1, fp); fclose(fp);
  }
  fg_mod_reg(acq, emf, fwi, xx, g);
  fwi->fcost = fwi->fcost_dat + fwi->fcost_mod;
  if (emf->mode == 1 && fwi->firstgrad) {
    s1 = 0; s2 = 0;
    for (k = 0; k < fwi->n; k++) { s1 += fabs(xx[k]); s2 += fabs(g[k]); }
    fwi->alpha = 2e-3*s1/s2;
  }
  for (k = 0; k < fwi->n; k++) g[k] *= fwi->alpha;
  if (fwi->firstgrad) fwi->firstgrad = 0;
  return fwi->fcost*fwi->alpha;
}

void Hv_fwi(float *x, float *v, float *Hv)
{
  (void)x; (void)v; (void)Hv;
}

int main(int argc, char **argv)
{
  acq_t *a;
  emf_t *e;
  setvbuf(stdout, NULL, _IOLBF, 0);
  initargs(argc, argv);
  a = calloc(1, sizeof(acq_t));
  e = calloc(1, sizeof(emf_t));
  if (!getparint("mode", &e->mode)) e->mode = 1;
  if (!getparint("verb", &e->verb)) e->verb = 1;
  if (e->mode == 0) err("fwi_runner runs mode=1 or mode=2");
  emf_init(e);
  acq_init(a, e);
  do_fwi(a, e);
  printf("fwi_runner done\n");
  return 0;
}
