/* emmi_cli.h -- host side of the libEMMI_MGFD command line, as the library
 * libemmi_cli.so exports it: parameter parsing, the survey and data files,
 * the data weighting, regularisation and preconditioning of the inversion.
 * The executable emmi_main runs the reference's own command lines (mode=0
 * modelling, mode=1 inversion, mode=2 gradient) with these functions around
 * the GPU solver of libemmi_mgfd.so.
 *
 * Every function keeps the name, argument meaning and error behaviour of the
 * reference function it replaces (cited per declaration); errors print
 * "\n<program>: <message>\n" on stderr and exit(1), as src/cstd.c:1531-1544
 * does.  The carriers emf_t/acq_t/fwi_t are those of emmi_mgfd.h.
 */
#ifndef EMMI_CLI_H
#define EMMI_CLI_H

#include "emmi_mgfd.h"

#ifdef __cplusplus
extern "C" {
#endif

/* rank of the process and number of processes: src/main.c:13 keeps them
 * for MPI; here they name the shot a call works on (shot_idx[iproc]) and
 * the number of shots of the run */
extern int iproc, nproc;

/* ------------------------------------------------------------------------- */
/* parameters (src/cstd.c:936-1432): key=value words of the command line   */
/* plus the key=value words of the file named by the last par=, the file's  */
/* first; the last occurrence of a key wins; lists are comma separated      */
/* ------------------------------------------------------------------------- */
void initargs(int argc, char **argv);                     /* src/cstd.c:936-950 */
int getparint(char *name, int *ptr);                       /* src/cstd.c:972-974 */
int getparfloat(char *name, float *ptr);                   /* src/cstd.c:996-998 */
int getpardouble(char *name, double *ptr);                 /* src/cstd.c:1000-1002 */
int getparstring(char *name, char **ptr);                  /* src/cstd.c:1004-1006 */
int getparstringarray(char *name, char **ptr);             /* src/cstd.c:1008-1010 */
int countparval(char *name);                               /* src/cstd.c:1224-1227 */
/* src/cstd.c:1531-1544 (err): message on stderr, exit(1) */
void emmi_cli_err(const char *fmt, ...);

/* ------------------------------------------------------------------------- */
/* survey and data files                                                     */
/* ------------------------------------------------------------------------- */
/* src/emf_init_free.c:11-134: frequencies, channels, model grid, the three
 * resistivity files (float32 [nz][ny][nx]), rho_water = smallest value */
void emf_init(emf_t *emf);
void emf_free(emf_t *emf);                                 /* src/emf_init_free.c:137-147 */
/* src/acq_init_free.c:13-262: fsrc/frec/fsrcrec tables, shots=, reciprocity,
 * domain checks; the shot is shot_idx[iproc] */
void acq_init(acq_t *acq, emf_t *emf);
void acq_free(acq_t *acq);                                 /* src/acq_init_free.c:265-280 */
/* src/read_write_data.c:14-35: one line per (channel, frequency, receiver) */
void write_data(acq_t *acq, emf_t *emf, char *fname, emmi_c64 ***dcal_fd);
/* src/read_write_data.c:39-76: emf_<shot>.txt into emf->dobs_fd */
void read_data(acq_t *acq, emf_t *emf);
/* src/read_write_data.c:80-102: rmse_<shot>.txt from emf->rmse */
void write_rmse(acq_t *acq, emf_t *emf);

/* ------------------------------------------------------------------------- */
/* data weighting                                                            */
/* ------------------------------------------------------------------------- */
/* src/frannor.c:90-140: N(0,1) floats by the Marsaglia-Tsang method as
 * Kahaner, Moler and Nash's RNOR states it; one state per process */
float frannor(void);
/* src/frannor.c:143-169: seed the state */
void srannor(int seed);
/* the state a fresh process starts from (src/frannor.c:77-85); a run of
 * several shots in one process resets it per shot, as each MPI rank of the
 * reference starts fresh */
void emmi_rannor_reset(void);
/* src/cal_uncertainty_noise.c:17-98: emf->delta_emf from the observed data,
 * and, when emf->addnoise, Gaussian noise of that size added to dobs_fd;
 * returns the sum of |noise|^2 */
float cal_uncertainty_noise(acq_t *acq, emf_t *emf);
/* src/fg_fwi.c:207-225: residuals beyond 5.5 skin depths of the source,
 * weighted twice by delta_emf and conjugated into emf->dres_fd, their
 * weighted moduli into emf->rmse; returns 0.5 * sum |weighted residual|^2 */
float emmi_data_misfit(acq_t *acq, emf_t *emf);

/* ------------------------------------------------------------------------- */
/* model regularisation and gradient preconditioning of the inversion        */
/* ------------------------------------------------------------------------- */
/* src/regularization.c:11-57: roughness c_h(|Dx x|^2 + |Dy x|^2) +
 * c_v |Dz x|^2 of an n1 x n2 x n3 volume (unit spacings, c_h = 1,
 * c_v = 0.03) and its gradient into g; d1..d3 are unused as there */
float regularization_tikhonov(float *x, float *g, int n1, int n2, int n3, float d1, float d2, float d3);
/* src/regularization.c:60-130: total variation sum sqrt(|grad x|^2 + 1e-14)
 * and its gradient into g */
float regularization_tv(float *x, float *g, int n1, int n2, int n3, float d1, float d2, float d3);
/* src/triangle_smoothing.c:63-107: triangle filters of half-widths r1, r2,
 * r3 along the three axes of mod[n3][n2][n1], repeated */
void triangle_smoothing(float ***mod, int n1, int n2, int n3, int r1, int r2, int r3, int repeat);
/* src/precondition.c:16-87: fwi->preco 1 divides g by the pseudo-Hessian
 * fwi->phess, 2 scales it by exp(depth below seabed / skin depth), 3 smooths
 * it (using fwi->grad as scratch) */
void precondition(emf_t *emf, fwi_t *fwi, float *g);
/* src/fg_mod_reg.c:19-94: adds the Tikhonov (fwi->gamma1) and TV
 * (fwi->gamma2) terms of x - fwi->xref below the seabed, cooled by
 * 0.8^fwi->iter, to g and to fwi->fcost_mod */
void fg_mod_reg(acq_t *acq, emf_t *emf, fwi_t *fwi, float *x, float *g);

#ifdef __cplusplus
}
#endif
#endif
