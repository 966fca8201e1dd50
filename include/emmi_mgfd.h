/* emmi_mgfd.h -- C-ABI of the MI355X multigrid Maxwell solve.
 *
 * Two layers:
 *  1. the reference's own entry points for this path, same names, argument
 *     meaning and ownership (the callers declare them locally at
 *     src/do_modelling.c:14-26 and src/fg_fwi.c:19-41 of libEMMI_MGFD), so a
 *     host program built against libEMMI_MGFD links this library instead;
 *  2. an extended device-resident API (emmi_*) with explicit handles, status
 *     codes, initial guesses and batches of right-hand sides.
 *
 * Layouts are the reference's: field vectors are complex double
 * [3][n3+1][n2+1][n1+1] with i fastest; cell arrays [n3][n2][n1]; input-grid
 * arrays [nz][ny][nx].
 */
#ifndef EMMI_MGFD_H
#define EMMI_MGFD_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
typedef struct { double re, im; } emmi_c128;   /* layout of double _Complex */
typedef struct { float re, im; } emmi_c64;     /* layout of float _Complex */
#else
typedef double _Complex emmi_c128;
typedef float _Complex emmi_c64;
#endif

/* ------------------------------------------------------------------------- */
/* carriers, layout-identical to include/emf.h:4-48, include/acq.h:4-19 and  */
/* include/fwi.h:4-28 of libEMMI_MGFD (3D/4D arrays are pointer tables over  */
/* one contiguous block, &a[0][0][0] being its base)                          */
/* ------------------------------------------------------------------------- */
#ifndef EMMI_NO_CARRIERS
typedef struct {
  int mode, verb, reciprocity;
  int nfreq;
  float *freqs, *omegas;
  float rho_air, rho_water, lextend;
  int nchsrc, nchrec;
  char **chsrc, **chrec;
  float x1min, x1max, x2min, x2max, x3min, x3max;
  float *xx, *yy, *zz;
  int n1, n2, n3;
  float d1min, d2min, d3min;
  float *x1, *x2, *x3;
  float *x1s, *x2s, *x3s;
  float *d1, *d2, *d3;
  float *d1s, *d2s, *d3s;
  float ***sigma11, ***sigma22, ***sigma33, ***invmur;
  emmi_c64 ***dcal_fd, ***dobs_fd, ***dres_fd;
  emmi_c64 ****E1, ****E2, ****E3;
  emmi_c64 ****H1, ****H2, ****H3;
  float ***mur, ***rho11, ***rho22, ***rho33;
  float **bathy;
  float rhomin, rhomax, depthmax;
  float ox, oy, oz;
  float dx, dy, dz;
  int nx, ny, nz, nb;
  int addnoise;
  float amp_perc, delta_phi, delta_theta;
  float noisefloorE, noisefloorH;
  float ***delta_emf;
  float ***rmse;
  float dsmute, drmute;
} emf_t;

typedef struct {
  int nsrc, nrec, nsubsrc, nsubrec;
  float lensrc, lenrec;
  float *src_x1, *src_x2, *src_x3, *src_azimuth, *src_dip;
  float *rec_x1, *rec_x2, *rec_x3, *rec_azimuth, *rec_dip;
  float x1min, x1max, x2min, x2max, x3min, x3max;
  int nsrc_total, nrec_total;
  int *shot_idx;
} acq_t;

typedef struct {
  int npar, n, preco, r1, r2, r3, repeat;
  int *idxpar;
  float *minpar, *maxpar;
  int iter, niter;
  float *xref;
  float gamma1, gamma2;
  int firstgrad;
  float alpha;
  float fcost_dat, fcost_mod, fcost;
  float *fcost_list;
  float ****grad, ****phess, ***Hv;
  float mse;
} fwi_t;
#endif

/* ------------------------------------------------------------------------- */
/* 1. drop-in entry points (reference semantics)                              */
/* ------------------------------------------------------------------------- */

/* replaces src/find_index.c:9-29: k with x[k] <= val < x[k+1] by bisection */
int find_index(int n, float *x, float val);
/* replaces src/create_nugrid.c:13-36: geometric grid, returns the ratio */
float create_nugrid(int n, float len, float dx, float *x);
/* replaces src/homogenization.c:15-111: overlap-weighted volume average */
void homogenization(float ***in, float ***out, int nx, int ny, int nz, float *xx, float *yy, float *zz,
                    int n1, int n2, int n3, float *x1, float *x2, float *x3);
/* replaces src/regrid_init_free.c:37-232 / :235-262: stretched MG grid around
 * the source and homogenised medium (bit-exact) */
void regrid_init(acq_t *acq, emf_t *emf);
void regrid_free(emf_t *emf);

/* replaces src/gmg.c:1704-1733: reads cycleopt, ncycle, v1, v2, tol,
 * isemicoarsen (and the extensions precision=, coarse_sweeps=) through
 * getpar; keeps emf (borrowed, must outlive gmg_apply) */
void gmg_init(emf_t *emf, int ifreq);
/* replaces src/gmg.c:1894-1932: solves from a zero guess on the GPU; x <- E,
 * b <- H; n must be 3(n1+1)(n2+1)(n3+1).  Host pointers. */
void gmg_apply(int n, emmi_c128 *b, emmi_c128 *x);
/* replaces src/gmg.c:1887-1890 */
void gmg_free(void);

/* replace src/inject_extract.c:26-120, :124-314, :317-474, :477-635,
 * :638-785 (host vectors, bit-exact) */
void inject_source(acq_t *acq, emf_t *emf, emmi_c128 *b, int ifreq);
void inject_adj_source(acq_t *acq, emf_t *emf, emmi_c128 *x, emmi_c128 *b, int ifreq);
void extract_emf_data(acq_t *acq, emf_t *emf, emmi_c128 *x, emmi_c128 *b, int ifreq);
void extract_electric_field(acq_t *acq, emf_t *emf, emmi_c128 *x, int ifreq);
void extract_magnetic_field(acq_t *acq, emf_t *emf, emmi_c128 *b, int ifreq);
/* replaces src/build_gradient.c:13-111 */
void build_gradient(acq_t *acq, emf_t *emf, fwi_t *fwi,
                    emmi_c64 ****E1f, emmi_c64 ****E2f, emmi_c64 ****E3f,
                    emmi_c64 ****E1a, emmi_c64 ****E2a, emmi_c64 ****E3a);

/* ------------------------------------------------------------------------- */
/* 2. extended API                                                            */
/* ------------------------------------------------------------------------- */
typedef struct {
  int cycleopt;       /* 1 V, 2 F (default), 3 W */
  int ncycle;         /* default 30 */
  int v1, v2;         /* default 1, 1 */
  int isemicoarsen;   /* default 1 */
  double tol;         /* relative residual ||f - A u|| / ||f||, default 1e-6 */
  int precision;      /* 0 complex FP64; 1 complex FP32 cycles in FP64 refinement */
  int coarse_sweeps;  /* line sweeps per coarsest-level smoothing, default 2 */
  int verb;
  int graphs;         /* 1 (default): replay each cycle as a captured HIP graph */
} emmi_params_t;

typedef struct {
  int cycles;         /* completed cycles */
  int status;         /* EMMI_CONVERGED, EMMI_NOT_CONVERGED, EMMI_NONFINITE */
  double rnorm0;      /* ||f|| (FP64 V- and W-cycles: the first in-cycle norm,
                         after the pre-smoothing of cycle 0, as src/gmg.c:1594-1599) */
  double rnorm;       /* last ||f - A u|| */
  int x0_rejected;    /* 1: the initial guess had a residual ||f - A x0|| not
                         below ||f|| (or not finite) and the solve started
                         from zero instead */
  double x0_rnorm;    /* ||f - A x0|| of the initial guess, 0 without one */
} emmi_stats_t;

enum { EMMI_CONVERGED = 0, EMMI_NOT_CONVERGED = 1, EMMI_NONFINITE = 2, EMMI_ERROR = -1 };

typedef struct emmi_solver emmi_solver_t;

void emmi_default_params(emmi_params_t *p);
const char *emmi_last_error(void);

/* A solver owns the level hierarchy of one (grid, medium, frequency).
 * x1..x3: float node coordinates (n+1 each, host); sig11..sig33: float
 * conductivities [n3][n2][n1], host pointers or device pointers when
 * sig_on_device != 0.  Returns NULL on error (see emmi_last_error). */
emmi_solver_t *emmi_solver_create(int n1, int n2, int n3, const float *x1, const float *x2, const float *x3,
                                  const float *sig11, const float *sig22, const float *sig33, int sig_on_device,
                                  float omega, const emmi_params_t *p);
/* Device vectors (complex double, 3(n1+1)(n2+1)(n3+1)): b source term
 * i omega mu0 J / vol as inject_source writes it, x <- E, h <- H (h may be b),
 * x0 initial guess or NULL.  Returns the status code. */
int emmi_solver_solve(emmi_solver_t *s, const void *b, void *x, void *h, const void *x0, emmi_stats_t *st);
/* (Solves first wait for all work queued on the device, so b and x0 may come
 * from kernels on any stream; they return when x and h are complete.) */
/* switch the frequency of a solver (the level hierarchy depends on the grid
 * and medium only, so the frequencies of one shot share it) */
int emmi_solver_set_omega(emmi_solver_t *s, float omega);
/* residual norms of the last solve, one per cycle; returns the count */
int emmi_solver_history(emmi_solver_t *s, double *out, int maxn);
void emmi_solver_destroy(emmi_solver_t *s);

/* Batched solves (no counterpart in the reference, which solves one
 * frequency per gmg_init/gmg_apply pair in the loop of
 * src/do_modelling.c:49-67).  A batch solver holds nslots right-hand sides
 * that share the grid and medium of one shot, each with its own frequency
 * omegas[k].  emmi_solver_solve_batch solves slots 0..n-1 together: every
 * kernel launch of a cycle serves all unconverged slots, and each slot stops
 * when its own residual test passes, so slot k gets exactly the iterates a
 * single solve at omegas[k] would produce.  b, x, h, x0 are tables of n
 * device vectors (x0 may be NULL or hold NULL entries); st receives n
 * records.  Returns the worst status of the batch or EMMI_ERROR. */
emmi_solver_t *emmi_solver_create_batch(int n1, int n2, int n3, const float *x1, const float *x2, const float *x3,
                                        const float *sig11, const float *sig22, const float *sig33,
                                        int sig_on_device, const float *omegas, int nslots, const emmi_params_t *p);
int emmi_solver_solve_batch(emmi_solver_t *s, int n, const void *const *b, void *const *x, void *const *h,
                            const void *const *x0, emmi_stats_t *st);
/* Survey solves (no counterpart in the reference, which runs one rank per
 * shot, src/acq_init_free.c:120-142): nmed grids and media of one multigrid
 * size -- x1[m]..x3[m], sig11[m]..sig33[m] as for emmi_solver_create, e.g.
 * the shots of a survey each regridded about its own source -- and nslots
 * <= 64 right-hand sides, slot k on grid slot_med[k] at omegas[k].  Every
 * kernel launch of a cycle serves all unconverged slots whatever their grid;
 * each slot stops at its own convergence test.  Solve with
 * emmi_solver_solve_batch. */
emmi_solver_t *emmi_solver_create_survey(int n1, int n2, int n3, int nmed, const float *const *x1,
                                         const float *const *x2, const float *const *x3, const float *const *sig11,
                                         const float *const *sig22, const float *const *sig33, int sig_on_device,
                                         const int *slot_med, const float *omegas, int nslots,
                                         const emmi_params_t *p);
int emmi_solver_set_omega_slot(emmi_solver_t *s, int slot, float omega);
int emmi_solver_history_slot(emmi_solver_t *s, int slot, double *out, int maxn);
int emmi_solver_slots(emmi_solver_t *s);
/* cycle limit of the following solves (the ncycle of the parameters) */
int emmi_solver_set_ncycle(emmi_solver_t *s, int ncycle);
/* No counterpart in the reference, whose solves always start at cycle 0
 * (src/gmg.c:1911-1926): cycle k of the following solves uses the
 * semi-coarsening pattern (first + k) % 3 instead of k % 3, so a solve that
 * continues from the iterate of `first` earlier cycles keeps the reference's
 * pattern sequence.  Default 0. */
int emmi_solver_set_first_cycle(emmi_solver_t *s, int first);
/* No counterpart in the reference: before a solve from guesses, blend each
 * guess with a multigrid iterate of the same right-hand side.  For slots
 * 0..n-1 (b: source terms as for emmi_solver_solve_batch; x1: iterates, e.g.
 * after one cycle from zero; xg: guesses, overwritten), xg <- x1 + beta
 * (xg - x1) with the complex beta that minimises ||f - A xg||, so the blend's
 * residual is never above x1's and the solver's rule never rejects it;
 * beta = 0 when the guess changes nothing or a norm is not finite.  beta
 * receives 2n doubles (re, im) or may be NULL.  Returns 0 or EMMI_ERROR. */
int emmi_solver_blend_guess(emmi_solver_t *s, int n, const void *const *b, const void *const *x1, void *const *xg,
                            double *beta);
/* Autotuner (no counterpart in the reference).  For every line-sweep launch
 * shape a solve of all the solver's slots issues -- precision, axis, line
 * length, lines per colour, right-hand sides per launch -- that the
 * process-wide table does not hold yet, times each substitution kernel
 * (one lane per line half; four lanes per line half with the right-hand
 * sides computed in the same launch; four lanes with a separate
 * right-hand-side launch) reps times on the slots' own vectors and records
 * the fastest; later launches of that shape use it (unless
 * emmi_set_line_kernel forces a kernel).  Overwrites the solver's vectors
 * (the next solve sets them) and drops its captured cycle graphs.  Returns
 * the number of shapes tuned or EMMI_ERROR.  The kernels differ in rounding
 * order, so solves of one right-hand side in different pack sizes may differ
 * in the last bits once the table holds entries. */
int emmi_solver_autotune(emmi_solver_t *s, int reps);
/* the table: rows[6k..6k+5] = value size (4 or 8), axis, line length (cells),
 * lines per colour, right-hand sides per launch, choice (0 one lane, 1 quad
 * fused, 2 quad separate); us[2k] = microseconds per launch of the choice,
 * us[2k+1] of the built-in rule.  Returns the number of entries (may exceed
 * maxn) */
int emmi_line_tuning(int *rows, double *us, int maxn);
void emmi_line_tuning_clear(void);
int emmi_solver_levels(emmi_solver_t *s, int *npatterns);
void emmi_solver_dims(emmi_solver_t *s, int pattern, int lev, int *n123);

/* Single-kernel checks on host vectors (complex double), used by the parity
 * tests: op 0 residual r = f - A u (in1 = u, in2 = f); op 1 one line
 * Gauss-Seidel sweep along `axis` (4 colours, reversed when iter is odd) of
 * u = in1 with f = in2; op 2 restriction of in1 from level lev to lev+1;
 * op 3 prolongation: out = in2 + P in1 (in1 on level lev+1, in2 on lev);
 * op 4 H = curl in1 / (i omega mu0) (level 0); op 5 coarsest-level smoothing
 * of u = in1 with f = in2; op 6 one vertex-patch Gauss-Seidel sweep (8 node
 * colours, reversed when iter is odd) of u = in1 with f = in2.  `precision`
 * selects the FP64 or FP32 kernels (the solver must have been created with
 * that precision). */
int emmi_solver_check_op(emmi_solver_t *s, int op, int pattern, int lev, int axis, int iter,
                         const emmi_c128 *in1, const emmi_c128 *in2, emmi_c128 *out);

/* Device-resident steps around the solve (stream may be NULL).  The source
 * term is computed with the arithmetic of inject_source and scattered into
 * the device vector b_dev; receiver data are gathered from the device E/H
 * vectors and written to emf->dcal_fd[.][ifreq][.] with the arithmetic of
 * extract_emf_data.  Both are bit-identical to the host entry points.
 * Return 0, or -1 with emmi_path_last_error() set. */
int emmi_inject_source_dev(acq_t *acq, emf_t *emf, int ifreq, void *b_dev, void *stream);
int emmi_extract_emf_data_dev(acq_t *acq, emf_t *emf, int ifreq, const void *x_dev, const void *h_dev,
                              void *stream);
/* inject_adj_source (src/inject_extract.c:124-314) on the device: the
 * adjoint source of the residuals emf->dres_fd[.][ifreq][.] into b_dev, the
 * magnetic dipole M of the H channels into x_dev (both complex double,
 * 3(n1+1)(n2+1)(n3+1)).  The scatters are accumulated on the host in the
 * reference's order, b += curl M runs on the device; bit-identical to the
 * host entry point. */
int emmi_inject_adj_source_dev(acq_t *acq, emf_t *emf, int ifreq, void *x_dev, void *b_dev, void *stream);
/* No counterpart in the reference: emmi_inject_adj_source_dev for callers
 * that cannot allocate device memory (code under HIP graph capture, serving
 * loops, harnesses that forbid hipMalloc).  Same arguments, same results bit
 * for bit; the device tables of the call (entry indices and values, dual
 * widths) go to a static device array of 2 MiB that each GPU holds for the
 * life of the process (enough for about 900 receivers with all six
 * channels).  A call whose tables exceed it returns -1 with an
 * emmi_path_last_error() naming the sizes.  It allocates no device memory;
 * calls that overlap on one GPU (other host threads, other streams) share
 * that GPU's array and race. */
int emmi_inject_adj_source_dev_noalloc(acq_t *acq, emf_t *emf, int ifreq, void *x_dev, void *b_dev, void *stream);
/* extract_electric_field (src/inject_extract.c:477-635) on the device: E of
 * the device field vector x_dev (complex double, 3(n1+1)(n2+1)(n3+1)) on the
 * model's cell centres, written to e_dev as complex float [3][nz][ny][nx]
 * (E1, E2, E3) instead of emf->E1..E3[ifreq]; the medium is read from emf
 * (sigma33 of the multigrid, rho33 of the model).  Same arithmetic as the
 * host entry point, bit for bit. */
int emmi_extract_electric_field_dev(acq_t *acq, emf_t *emf, int ifreq, const void *x_dev, void *e_dev,
                                    void *stream);
/* extract_magnetic_field (src/inject_extract.c:638-785) on the device: H of
 * the device field vector h_dev (complex double, as gmg_apply returns it in
 * b) on the model's cell centres, written to e_dev as complex float
 * [3][nz][ny][nx] (H1, H2, H3) instead of emf->H1..H3[ifreq].  Same
 * arithmetic as the host entry point, bit for bit. */
int emmi_extract_magnetic_field_dev(acq_t *acq, emf_t *emf, int ifreq, const void *h_dev, void *e_dev,
                                    void *stream);
/* build_gradient (src/build_gradient.c:13-111) on the device: ef_dev[f] and
 * ea_dev[f] (f < emf->nfreq <= 32) are the forward and adjoint fields of
 * frequency f as emmi_extract_electric_field_dev writes them; grad_dev and,
 * when fwi->preco == 1, phess_dev receive [npar][nz][ny][nx] float
 * (npar <= 8) instead of fwi->grad / fwi->phess, mutes included.  The
 * gradient is bit-identical to the host entry point; the pseudo-Hessian uses
 * the device hypot and agrees to float rounding. */
int emmi_build_gradient_dev(acq_t *acq, emf_t *emf, fwi_t *fwi, const void *const *ef_dev, const void *const *ea_dev,
                            float *grad_dev, float *phess_dev, void *stream);
const char *emmi_path_last_error(void);

/* ------------------------------------------------------------------------- */
/* 3. initial-guess network kernels (no counterpart in the reference)        */
/* ------------------------------------------------------------------------- */
/* A 3D U-Net predicts an initial field x0 for the solve from the medium, the
 * source position and the frequency.  Tensors are channel-last
 * [Z][Y][X][C], C a multiple of 16, BF16 (16-bit storage) unless noted;
 * accumulation is FP32.  Convolutions are 3x3x3, zero padding 1, implicit
 * GEMM on the matrix cores.  Weights: FP32 master [Cout][27*Cin]
 * (tap = 9 dz + 3 dy + dx), converted by emmi_x0_prep_weights to the BF16
 * forward layout [Cout][kpad(Cin)] and the flipped, transposed layout
 * [Cin][kpad(Cout)] used for the input gradient.  Epilogues (epi): 0 bias +
 * LeakyReLU(slope) -> BF16 [V][Cout]; 1 bias, FP32 [V][cout_store]; 2 input
 * gradient times the LeakyReLU derivative of act -> BF16; 3 plain BF16.
 * All return 0 or -1; stream may be NULL. */
int emmi_x0_kpad(int cin);
int emmi_x0_conv_fwd(const void *x, int Z, int Y, int X, int Cin, const void *w, int Cout, const float *bias, int epi,
                     float slope, const void *act, void *out_bf16, float *out_f32, int cout_store, void *stream);
/* weight gradient [Cout][27*Cin] (and bias gradient, db may be NULL) of a
 * convolution with input x [V][Cin] and output gradient dy [V][Cout];
 * part: scratch of ceil(Z/zslab)*27*Cin*Cout floats; Cin and Cout multiples
 * of 16, Cout at most 256 */
int emmi_x0_conv_dw(const void *x, const void *dy, int Z, int Y, int X, int Cin, int Cout, float *part, int zslab,
                    float *dw, float *db, void *stream);
int emmi_x0_prep_weights(const float *master, int Cin, int Cout, void *w, void *wt, void *stream);
int emmi_x0_pool2(const void *in, int Z, int Y, int X, int C, void *out, void *stream);
int emmi_x0_pool2_bwd(const void *dout, int Z, int Y, int X, int C, void *din, int accumulate, void *stream);
/* out [Z][Y][X][C1+C2] = concat(nearest 2x upsampling of in, skip) and its backward */
int emmi_x0_up2_cat(const void *in, const void *skip, int Z, int Y, int X, int C1, int C2, void *out, void *stream);
int emmi_x0_up2_cat_bwd(const void *dout, int Z, int Y, int X, int C1, int C2, void *din, void *dskip, int accumulate,
                        void *stream);
int emmi_x0_dact(void *g, const void *a, long n, float slope, void *stream);
int emmi_x0_mse(const float *out, const float *tgt, const float *mask, long nv, int cs, float inv_count, void *grad,
                float *part, void *stream);
int emmi_x0_adam(float *p, const float *g, float *m1, float *m2, long n, float lr, float b1, float b2, float eps,
                 int step, float clip, void *stream);
int emmi_x0_pack_bf16(const float *in, long nv, int C, int Cp, void *out, void *stream);
/* The network refines x1, the field after one multigrid cycle: per field
 * component it outputs dL and a relative phasor (1+q1, q2), and
 * x0_a = x1_a exp(dL) (1+q1 + i q2)/|1+q1 + i q2| (zero outputs give x1).
 * out_to_edges builds x0 from the output [n3][n2][n1][9]; edges_to_target
 * builds the training targets from a solved field E and x1, with loss
 * weights 1 (dL) and wp (phasor) where |E| and |x1| exceed eref. */
int emmi_x0_out_to_edges(const float *o, const void *x1, int n1, int n2, int n3, void *x0, void *stream);
/* the same with each edge's correction averaged over the four cells that
 * share the edge; cells with s11 < sig_air (float conductivities [n3][n2][n1],
 * device, may be NULL) take no correction */
int emmi_x0_out_to_edges_avg(const float *o, const void *x1, const float *s11, float sig_air, int n1, int n2, int n3,
                             void *x0, void *stream);
int emmi_x0_edges_to_target(const void *E, const void *x1, int n1, int n2, int n3, double eref, float wp, float *tgt,
                            float *mask, void *stream);
/* input features [n1*n2*n3][32] from the medium, the source offset in skin
 * depths and, when x1 is given, the field after one multigrid cycle */
int emmi_x0_features(const float *s11, const float *s33, const float *xs1, const float *xs2, const float *xs3,
                     const float *d1, const float *d2, const float *d3, int n1, int n2, int n3, float sx, float sy,
                     float sz, float delta, float hsrc, const void *x1, double eref, void *feat, void *stream);
int emmi_x0_scale(void *v, long n, double s, void *stream);
/* fill nwords 32-bit words with value */
int emmi_x0_fill(void *p, long nwords, unsigned value, void *stream);

/* device memory helpers for hosts without a GPU runtime of their own */
void *emmi_device_alloc(size_t bytes);
void emmi_device_free(void *p);
int emmi_memcpy_h2d(void *dst, const void *src, size_t bytes);
int emmi_memcpy_d2h(void *dst, const void *src, size_t bytes);
int emmi_memcpy_d2d(void *dst, const void *src, size_t bytes);
int emmi_device_sync(void);
/* Line-substitution kernel of the following launches (no counterpart in the
 * reference; for tests and tuning): 0 by line count (default), 1 one lane per
 * line half, 2 four lanes per line half, 3 four lanes with the right-hand
 * sides built by the substitution launch itself, 4 four lanes with the
 * right-hand sides built by a launch of their own.  All solve the same line
 * systems exactly.  Graphs a solver captured earlier keep the kernels they
 * captured.  Returns the previous mode, or -1 for an unknown mode. */
int emmi_set_line_kernel(int mode);
/* kernel timing: accumulated milliseconds, launch counts and compulsory
 * bytes (each input read once, each output written once per launch at its
 * level's shape; a full line sweep, all four colours, counts 3 Ne s + 3 Nc c)
 * per kernel class (0 residual, 1 line sweep, 2 restriction, 3 prolongation,
 * 4 other, 5 line right-hand sides), measured with HIP events recorded on
 * the stream of each launch (by the dispatch itself for the line kernels)
 * while enabled; enabling clears the counters.  emmi_timing_read_layout
 * gives, per class, the bytes of the kernels' own data layout (for the line
 * kernels: their stored factor and line records as well). */
/* Kernel rates on the solver's own grid (no counterpart in the reference):
 * the level-0 kernels of slot 0, in the solver's cycle precision, each
 * launched reps times back to back on the solver's stream between HIP
 * events; ms[k] is the average launch time, bytes[k] its compulsory bytes,
 * k = 0 residual (+ norm partials), 1 restriction and 2 prolongation
 * between level 0 and level 1 of semi-coarsening pattern 0, 3..5 a full line
 * sweep (four colours, right-hand sides included) along x, y, z.  Overwrites
 * slot 0's level vectors.  Returns 6 or EMMI_ERROR; n >= 6. */
int emmi_solver_kernel_rates(emmi_solver_t *s, int reps, double *ms, double *bytes, int n);
void emmi_timing_enable(int on);
int emmi_timing_read(double *ms, long *count, double *bytes, int n);
int emmi_timing_read_layout(double *bytes, int n);

#ifdef __cplusplus
}
#endif
#endif
