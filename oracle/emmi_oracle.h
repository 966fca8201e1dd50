/* emmi_oracle.h -- CPU restatement of libEMMI_MGFD's multigrid Maxwell path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker the HIP path is
 * compared against; only tests/, __graft_entry__.smoke() and bench.py's CPU
 * baseline leg may load it.  The product library never links or calls it.
 *
 * Arrays are flat and contiguous with the reference's layouts:
 *   cell arrays     [n3][n2][n1]               (i fastest)
 *   edge vectors    [3][n3+1][n2+1][n1+1]      complex double, (re,im) pairs
 * Grid arrays (x, xs, d, ds) have n+1 entries.
 */
#ifndef EMMI_ORACLE_H
#define EMMI_ORACLE_H

#include <complex.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef double _Complex ocplx;

/* ---- grid construction (bit-exact contract) ---------------------------- */
int   emo_find_index(int n, const float *x, float val);
float emo_create_nugrid(int n, float len, float dx, float *x);
void  emo_staggered(int n, const float *x, float *xs, float *d, float *ds);
void  emo_homogenize(const float *in, float *out, int nx, int ny, int nz,
                     const float *xx, const float *yy, const float *zz,
                     int n1, int n2, int n3,
                     const float *x1, const float *x2, const float *x3);

typedef struct {
  int nx, ny, nz, nb;
  float dx, dy, dz, ox, oy, oz;
  float lextend, rho_air;
  int n1, n2, n3;
  float d1min, d2min, d3min;
  const float *rho11, *rho22, *rho33;     /* [nz][ny][nx] */
} emo_model_t;

/* Outputs of the regridding, caller-allocated:
 *   xx[nx+2nb+1], yy[..], zz[..]; x1,x1s,d1,d1s[n1+1] (same for 2,3);
 *   sigma11, sigma22, sigma33 [n3][n2][n1]. */
typedef struct {
  float *xx, *yy, *zz;
  float *x1, *x2, *x3, *x1s, *x2s, *x3s, *d1, *d2, *d3, *d1s, *d2s, *d3s;
  float *sigma11, *sigma22, *sigma33;
} emo_grid_t;

void emo_regrid(const emo_model_t *m, float src_x, float src_y, float src_z, emo_grid_t *g);

/* ---- multigrid levels ---------------------------------------------------- */
typedef struct emo_level emo_level_t;

/* Level 0 from the float MG grid and float conductivities; mu = 1*volume. */
emo_level_t *emo_level0_new(int n1, int n2, int n3,
                            const float *x1, const float *x2, const float *x3,
                            const float *sig11, const float *sig22, const float *sig33);
/* Coarse level for semi-coarsening pattern `icycle` (axis icycle%3 kept when
 * isemicoarsen), restricting the medium of `fine`. */
emo_level_t *emo_level_coarsen(const emo_level_t *fine, int icycle, int isemicoarsen);
void emo_level_free(emo_level_t *l);
void emo_level_dims(const emo_level_t *l, int *n123, int *sc);
/* access to the level's double arrays: which 0..2 = sigma11..33, 3 = mu */
const double *emo_level_cells(const emo_level_t *l, int which);
const double *emo_level_axis(const emo_level_t *l, int axis, int kind); /* kind 0 x, 1 xs, 2 d, 3 ds */

/* r = f - A u (interior edges), r = f elsewhere */
void emo_residual(const emo_level_t *l, double omega_mu0, const ocplx *u, const ocplx *f, ocplx *r);
/* y = A u on interior edges, 0 elsewhere */
void emo_apply(const emo_level_t *l, double omega_mu0, const ocplx *u, ocplx *y);
/* line block Gauss-Seidel along `axis`; line order lexicographic (forward
 * if iter even), or by colour when colour_order != 0: the four classes
 * (pb%2, pc%2) in sequence, each class in lexicographic order. */
void emo_line_gs(const emo_level_t *l, double omega_mu0, int axis, int iter, int colour_order,
                 ocplx *u, const ocplx *f);
/* one line solve only (for element-wise kernel checks) */
void emo_line_solve(const emo_level_t *l, double omega_mu0, int axis, int pb, int pc,
                    ocplx *u, const ocplx *f);
/* point (vertex patch) Gauss-Seidel; colour_order != 0 sweeps the eight
 * classes (i%2, j%2, k%2) in sequence */
void emo_point_gs(const emo_level_t *l, double omega_mu0, int iter, int colour_order,
                  ocplx *u, const ocplx *f);
void emo_restrict(const emo_level_t *fine, const emo_level_t *coarse, const ocplx *rf, ocplx *fc);
void emo_prolong(const emo_level_t *fine, const emo_level_t *coarse, const ocplx *uc, ocplx *uf);
void emo_curl_h(const emo_level_t *l, double omega_mu0, const ocplx *u, ocplx *h);

/* ---- full solve ---------------------------------------------------------- */
typedef struct {
  int cycleopt;      /* 1 V, 2 F, 3 W */
  int ncycle, v1, v2, isemicoarsen;
  double tol;
  int colour_order;  /* 0: reference lexicographic sweeps */
  int verb;
} emo_params_t;

void emo_default_params(emo_params_t *p);
/* Solve A u = b * dual volume from u = 0 (src/gmg.c:1894-1932).  On return
 * x = E and b = H.  hist (length >= ncycle) receives the level-0 residual
 * norm per cycle; returns the number of norms written. */
int emo_gmg_solve(int n1, int n2, int n3, const float *x1, const float *x2, const float *x3,
                  const float *sig11, const float *sig22, const float *sig33, float omega,
                  const emo_params_t *p, ocplx *b, ocplx *x, double *hist);

/* ---- sources, receivers, gradient ---------------------------------------- */
typedef struct {
  int n1, n2, n3;
  const float *x1, *x2, *x3, *x1s, *x2s, *x3s, *d1, *d2, *d3;
  const float *sigma33;          /* MG grid [n3][n2][n1] */
  float rho_water;
} emo_mg_t;

/* channel codes: 0 Ex 1 Ey 2 Ez 3 Hx 4 Hy 5 Hz */
void emo_inject_source(const emo_mg_t *g, int nsrc, const float *sx, const float *sy, const float *sz,
                       const float *az, const float *dip, int nch, const int *ch, float omega, ocplx *b);
void emo_inject_adj_source(const emo_mg_t *g, int nrec, const float *rx, const float *ry, const float *rz,
                           const float *az, const float *dip, int nch, const int *ch,
                           const float _Complex *dres /* [nch][nrec] for this frequency */,
                           float omega, ocplx *x, ocplx *b);
void emo_extract_data(const emo_mg_t *g, int nrec, const float *rx, const float *ry, const float *rz,
                      const float *az, const float *dip, int nch, const int *ch,
                      const ocplx *x, const ocplx *b, float _Complex *dcal /* [nch][nrec] */);
/* input-grid cell-centre interpolation; E3 uses rho33 of the input model */
void emo_extract_efield(const emo_mg_t *g, int nx, int ny, int nz, float ox, float oy, float oz,
                        float dx, float dy, float dz, const float *rho33, const ocplx *x,
                        float _Complex *E1, float _Complex *E2, float _Complex *E3);
void emo_extract_hfield(const emo_mg_t *g, int nx, int ny, int nz, float ox, float oy, float oz,
                        float dx, float dy, float dz, const ocplx *b,
                        float _Complex *H1, float _Complex *H2, float _Complex *H3);

typedef struct {
  int nx, ny, nz;
  float ox, oy, oz, dx, dy, dz;
  const float *rho11, *rho33;   /* [nz][ny][nx] */
  const float *bathy;           /* [ny][nx] */
  float dsmute, drmute;
  int npar;
  const int *idxpar;            /* 1 = rho_h, 2 = rho_v */
  int preco;
} emo_gradspec_t;

void emo_build_gradient(const emo_gradspec_t *s, int nfreq,
                        const float _Complex *E1f, const float _Complex *E2f, const float _Complex *E3f,
                        const float _Complex *E1a, const float _Complex *E2a, const float _Complex *E3a,
                        int nsrc, const float *sx, const float *sy, const float *sz,
                        int nrec, const float *rx, const float *ry, const float *rz,
                        float *grad, float *phess);

#ifdef __cplusplus
}
#endif
#endif
