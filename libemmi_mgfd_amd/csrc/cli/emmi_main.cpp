// emmi_main.cpp -- the libEMMI_MGFD program (src/main.c:24-84) on the GPU
// solver: the same command lines, parameter files, input files and output
// files, with every solve of a shot run as one batch on the device.
//
//   mode=0  modelling (src/do_modelling.c:28-104): emf_<shot>.txt, and with
//           icheck=1 the fields Ex.bin .. Hz.bin
//   mode=1  inversion (src/do_fwi.c:43-240): l-BFGS with the bisection
//           line search, regularisation and preconditioning around the
//           misfit and gradient evaluations of src/fg_fwi.c:139-305;
//           iterate.txt, rmse_misfit.txt, Rv/Rh_iteration_NN and the files
//           of each evaluation
//   mode=2  one misfit and gradient evaluation (src/do_fwi.c:43-150,
//           src/fg_fwi.c:43-305): obs_/syn_/rmse_<shot>.txt,
//           gradient_dat_Rv/_Rh and the regularisation gradients
//   method=1 (Gauss-Newton) is refused: its Hessian-vector product is empty
//           in the reference (src/fg_fwi.c:307-311).
//
// The reference runs one MPI rank per shot.  This program runs the ranks of
// such a job one after another in one process on one GPU: shots=i,j,k (as
// for the reference) or nproc=N (shots 1..N, the reference's mpirun -n N);
// rank r works on shot shot_idx[r] exactly as the reference's rank r would,
// its noise generator starting from the fresh state, and the sums the
// reference forms with MPI_Allreduce (noise energy, misfit, gradient) are
// formed in rank order.
//
// Extensions (parameters the reference does not read): precision=0|1 (FP64
// cycles, or FP32 cycles inside FP64 refinement, default 1; either stops at
// the same FP64 residual test), graphs=0|1, coarse_sweeps=N, nproc=N,
// batch_shots=N (modes 1 and 2: the shots solved together in one survey
// solve; default 24 / nfreq), autotune=0|1 (time the line-sweep kernels on
// each new launch shape and use the fastest, default 0).
#include <cmath>
#include <cstdio>
#include <cstring>
#include <ctime>
#include <thread>
#include <vector>

#include "arrays.h"
#include "emmi_cli.h"
#include "inversion.h"

using namespace emmi_cli;

namespace {

int par_int(const char* k, int def) {
  int v = def;
  return getparint(const_cast<char*>(k), &v) ? v : def;
}
float par_float(const char* k, float def) {
  float v = def;
  return getparfloat(const_cast<char*>(k), &v) ? v : def;
}

void check(int rc, const char* what) {
  if (rc < 0) emmi_cli_err("%s: %s", what, emmi_path_last_error());
}

// the multigrid parameters of src/gmg.c:1712-1717 plus the extensions
emmi_params_t solver_params(int verb) {
  emmi_params_t p;
  emmi_default_params(&p);
  p.cycleopt = par_int("cycleopt", 2);
  p.ncycle = par_int("ncycle", 30);
  p.v1 = par_int("v1", 1);
  p.v2 = par_int("v2", 1);
  double tol = 1e-6;
  if (getpardouble(const_cast<char*>("tol"), &tol)) p.tol = tol; else p.tol = 1e-6;
  p.isemicoarsen = par_int("isemicoarsen", 1);
  p.precision = par_int("precision", 1);
  p.coarse_sweeps = par_int("coarse_sweeps", p.coarse_sweeps);
  p.graphs = par_int("graphs", 1);
  p.verb = verb;
  return p;
}

// device state of one shot: a batch solver over the shot's multigrid grid,
// one slot per frequency, with the slot vectors b (source in, H out) and x
// (E out); built after regrid_init, released before regrid_free
struct ShotSolve {
  emf_t* emf;
  long nv = 0;
  emmi_solver_t* solver = nullptr;
  std::vector<void*> b, x;
  std::vector<emmi_stats_t> st;

  ShotSolve(emf_t* e, const emmi_params_t& p) : emf(e) {
    nv = 3L * (e->n1 + 1) * (e->n2 + 1) * (e->n3 + 1);
    solver = emmi_solver_create_batch(e->n1, e->n2, e->n3, e->x1, e->x2, e->x3, &e->sigma11[0][0][0],
                                      &e->sigma22[0][0][0], &e->sigma33[0][0][0], 0, e->omegas, e->nfreq, &p);
    if (!solver) emmi_cli_err("solver: %s", emmi_last_error());
    for (int f = 0; f < e->nfreq; ++f) {
      b.push_back(emmi_device_alloc(sizeof(emmi_c128) * nv));
      x.push_back(emmi_device_alloc(sizeof(emmi_c128) * nv));
      if (!b.back() || !x.back()) emmi_cli_err("device allocation of the field vectors failed");
    }
    st.resize(e->nfreq);
  }
  ~ShotSolve() {
    for (void* p : b) emmi_device_free(p);
    for (void* p : x) emmi_device_free(p);
    emmi_solver_destroy(solver);
  }
  // all frequencies at once from zero initial guesses (src/gmg.c:1900);
  // H overwrites the source vector, as gmg_apply does (src/gmg.c:1929)
  void solve(const char* what) {
    const int rc = emmi_solver_solve_batch(solver, emf->nfreq, const_cast<const void* const*>(b.data()), x.data(),
                                           b.data(), nullptr, st.data());
    if (rc == EMMI_ERROR) emmi_cli_err("%s solve: %s", what, emmi_last_error());
    for (int f = 0; f < emf->nfreq; ++f) {
      if (emf->verb)
        printf("%s solve freq=%g Hz: %d cycles, rnorm/rnorm0=%e%s\n", what, emf->freqs[f], st[f].cycles,
               st[f].rnorm0 > 0 ? st[f].rnorm / st[f].rnorm0 : 0.0,
               st[f].status == EMMI_CONVERGED ? "" : " (not converged)");
      if (st[f].status == EMMI_NONFINITE) emmi_cli_err("%s solve of freq=%g Hz produced non-finite values", what,
                                                       emf->freqs[f]);
    }
  }
};

// The shots of a gradient evaluation solved together: each shot keeps its own
// grid and medium (a copy of emf after its regrid_init, with the rank's data
// tables bound), and one survey solver holds all of them with one slot per
// (shot, frequency), slot j*nfreq + f.  Every kernel launch of a cycle serves
// every unconverged slot; each slot stops at its own convergence test, so it
// gets the iterates a solve of that shot alone would give.
struct ShotGroup {
  int nf = 0, ns = 0;
  long nv = 0;
  std::vector<emf_t> E;
  std::vector<acq_t*> acq;
  emmi_solver_t* solver = nullptr;
  std::vector<void*> b, x;
  std::vector<emmi_stats_t> st;

  // base: the emf of the job (model, frequencies, parameters); emfs[j] is
  // base with rank j's data tables and verbosity
  ShotGroup(const std::vector<emf_t>& emfs, const std::vector<acq_t*>& acqs, const emmi_params_t& p)
      : E(emfs), acq(acqs) {
    ns = (int)E.size();
    nf = E[0].nfreq;
    // the shots' regrids are independent host work (no shared state)
    std::vector<std::thread> th;
    for (int j = 0; j < ns; ++j) th.emplace_back([this, j] { regrid_init(acq[j], &E[j]); });
    for (auto& t : th) t.join();
    const emf_t& e = E[0];
    nv = 3L * (e.n1 + 1) * (e.n2 + 1) * (e.n3 + 1);
    std::vector<const float*> x1(ns), x2(ns), x3(ns), s11(ns), s22(ns), s33(ns);
    for (int j = 0; j < ns; ++j) {
      x1[j] = E[j].x1; x2[j] = E[j].x2; x3[j] = E[j].x3;
      s11[j] = &E[j].sigma11[0][0][0]; s22[j] = &E[j].sigma22[0][0][0]; s33[j] = &E[j].sigma33[0][0][0];
    }
    std::vector<int> med(ns * nf);
    std::vector<float> om(ns * nf);
    for (int j = 0; j < ns; ++j)
      for (int f = 0; f < nf; ++f) {
        med[j * nf + f] = j;
        om[j * nf + f] = E[j].omegas[f];
      }
    solver = emmi_solver_create_survey(e.n1, e.n2, e.n3, ns, x1.data(), x2.data(), x3.data(), s11.data(),
                                       s22.data(), s33.data(), 0, med.data(), om.data(), ns * nf, &p);
    if (!solver) emmi_cli_err("solver: %s", emmi_last_error());
    if (par_int("autotune", 0) && emmi_solver_autotune(solver, 5) < 0)
      emmi_cli_err("autotune: %s", emmi_last_error());
    for (int k = 0; k < ns * nf; ++k) {
      b.push_back(emmi_device_alloc(sizeof(emmi_c128) * nv));
      x.push_back(emmi_device_alloc(sizeof(emmi_c128) * nv));
      if (!b.back() || !x.back()) emmi_cli_err("device allocation of the field vectors failed");
    }
    st.resize(ns * nf);
  }
  ~ShotGroup() {
    for (void* q : b) emmi_device_free(q);
    for (void* q : x) emmi_device_free(q);
    emmi_solver_destroy(solver);
    for (int j = 0; j < ns; ++j) regrid_free(&E[j]);
  }
  void* bs(int j, int f) { return b[j * nf + f]; }
  void* xs(int j, int f) { return x[j * nf + f]; }
  // all slots from zero initial guesses (src/gmg.c:1900); H overwrites the
  // source vector, as gmg_apply does (src/gmg.c:1929)
  void solve(const char* what) {
    const int rc = emmi_solver_solve_batch(solver, ns * nf, const_cast<const void* const*>(b.data()), x.data(),
                                           b.data(), nullptr, st.data());
    if (rc == EMMI_ERROR) emmi_cli_err("%s solve: %s", what, emmi_last_error());
    for (int j = 0; j < ns; ++j)
      for (int f = 0; f < nf; ++f) {
        const emmi_stats_t& t = st[j * nf + f];
        if (E[j].verb)
          printf("%s solve freq=%g Hz: %d cycles, rnorm/rnorm0=%e%s\n", what, E[j].freqs[f], t.cycles,
                 t.rnorm0 > 0 ? t.rnorm / t.rnorm0 : 0.0, t.status == EMMI_CONVERGED ? "" : " (not converged)");
        else if (t.status == EMMI_NOT_CONVERGED)
          printf("%s solve of shot %d, freq=%g Hz: not converged after %d cycles\n", what, acq[j]->shot_idx[0],
                 E[j].freqs[f], t.cycles);
        if (t.status == EMMI_NONFINITE)
          emmi_cli_err("%s solve of freq=%g Hz produced non-finite values", what, E[j].freqs[f]);
      }
  }
};

void write_bin(const char* name, const void* p, size_t bytes) {
  FILE* fp = fopen(name, "wb");
  if (!fp) emmi_cli_err("cannot open %s for writing", name);
  if (fwrite(p, bytes, 1, fp) != 1) emmi_cli_err("write of %s failed", name);
  fclose(fp);
}

// ---------------------------------------------------------------------------
// mode 0: src/do_modelling.c:28-104 for one rank
// ---------------------------------------------------------------------------
void do_modelling(acq_t* acq, emf_t* emf, const emmi_params_t& p) {
  const int icheck = par_int("icheck", 0);
  const size_t nx = emf->nx, ny = emf->ny, nz = emf->nz, nf = emf->nfreq;
  if (icheck) {
    emf->E1 = alloc4<emmi_c64>(nx, ny, nz, nf);
    emf->E2 = alloc4<emmi_c64>(nx, ny, nz, nf);
    emf->E3 = alloc4<emmi_c64>(nx, ny, nz, nf);
    emf->H1 = alloc4<emmi_c64>(nx, ny, nz, nf);
    emf->H2 = alloc4<emmi_c64>(nx, ny, nz, nf);
    emf->H3 = alloc4<emmi_c64>(nx, ny, nz, nf);
  }
  emf->dcal_fd = alloc3<emmi_c64>(acq->nrec, nf, emf->nchrec);
  regrid_init(acq, emf);
  {
    ShotSolve S(emf, p);
    for (int f = 0; f < emf->nfreq; ++f) check(emmi_inject_source_dev(acq, emf, f, S.b[f], nullptr), "inject_source");
    S.solve("forward");
    for (int f = 0; f < emf->nfreq; ++f)
      check(emmi_extract_emf_data_dev(acq, emf, f, S.x[f], S.b[f], nullptr), "extract_emf_data");
    if (icheck) {
      std::vector<emmi_c128> hx(S.nv), hb(S.nv);
      for (int f = 0; f < emf->nfreq; ++f) {
        if (emmi_memcpy_d2h(hx.data(), S.x[f], sizeof(emmi_c128) * S.nv) ||
            emmi_memcpy_d2h(hb.data(), S.b[f], sizeof(emmi_c128) * S.nv))
          emmi_cli_err("copy of the fields to the host failed");
        extract_electric_field(acq, emf, hx.data(), f);
        extract_magnetic_field(acq, emf, hb.data(), f);
      }
    }
  }
  regrid_free(emf);
  char fname[32];
  snprintf(fname, sizeof fname, "emf_%04d.txt", acq->shot_idx[iproc]);
  write_data(acq, emf, fname, emf->dcal_fd);
  if (icheck) {
    if (emf->verb) {
      const size_t bytes = sizeof(emmi_c64) * nf * nx * ny * nz;
      write_bin("Ex.bin", &emf->E1[0][0][0][0], bytes);
      write_bin("Ey.bin", &emf->E2[0][0][0][0], bytes);
      write_bin("Ez.bin", &emf->E3[0][0][0][0], bytes);
      write_bin("Hx.bin", &emf->H1[0][0][0][0], bytes);
      write_bin("Hy.bin", &emf->H2[0][0][0][0], bytes);
      write_bin("Hz.bin", &emf->H3[0][0][0][0], bytes);
    }
    emmi_c64**** f6[] = {emf->E1, emf->E2, emf->E3, emf->H1, emf->H2, emf->H3};
    for (auto a : f6) free4(a);
    emf->E1 = emf->E2 = emf->E3 = emf->H1 = emf->H2 = emf->H3 = nullptr;
  }
  free3(emf->dcal_fd);
  emf->dcal_fd = nullptr;
}

// ---------------------------------------------------------------------------
// modes 1 and 2: misfit and gradient evaluations (src/fg_fwi.c) inside the
// inversion loop (src/do_fwi.c)
// ---------------------------------------------------------------------------
struct Rank {
  acq_t acq{};
  emmi_c64 ***dobs = nullptr, ***dcal = nullptr, ***dres = nullptr;
  float ***delta = nullptr, ***rmse = nullptr;
};

class Inversion {
 public:
  // src/fg_fwi.c:43-137 (fg_fwi_init) for every rank, with the parameters
  // src/do_fwi.c:55-100 reads
  Inversion(emf_t* emf, const emmi_params_t& p, int verb_set, int verb_par)
      : emf_(emf), p_(p), verb_set_(verb_set), verb_par_(verb_par) {
    char* fbathy;
    if (!getparstring(const_cast<char*>("fbathy"), &fbathy)) emmi_cli_err("Need fbathy= for CSEM inversion");
    emf->addnoise = par_int("addnoise", 1);
    emf->amp_perc = par_float("amp_perc", 0.03f);
    emf->delta_phi = par_float("delta_phi", 0);
    emf->delta_theta = par_float("delta_theta", 0);
    emf->noisefloorE = par_float("noisefloorE", 1e-15f);
    emf->noisefloorH = par_float("noisefloorH", 1e-13f);
    emf->dsmute = par_float("dsmute", 800);
    emf->drmute = par_float("drmute", 200);
    fwi_.gamma1 = par_float("gamma1", 10.);
    fwi_.gamma2 = par_float("gamma2", 0.);
    fwi_.r1 = par_int("r1", 3);
    fwi_.r2 = par_int("r2", 3);
    fwi_.r3 = par_int("r3", 1);
    fwi_.repeat = par_int("repeat", 1);
    fwi_.preco = par_int("preco", 1);
    fwi_.npar = par_int("npar", 2);
    if (fwi_.npar < 1 || fwi_.npar > 8) emmi_cli_err("npar=%d outside 1..8", fwi_.npar);
    // the reference leaves idxpar unset when idxpar= is absent; here it
    // defaults to 1..npar (rho_h, then rho_v)
    idxpar_.resize(fwi_.npar);
    for (int j = 0; j < fwi_.npar; ++j) idxpar_[j] = j + 1;
    if (countparval(const_cast<char*>("idxpar")) > fwi_.npar) emmi_cli_err("idxpar= has more than npar=%d values", fwi_.npar);
    getparint(const_cast<char*>("idxpar"), idxpar_.data());
    fwi_.idxpar = idxpar_.data();
    nxyz_ = static_cast<size_t>(emf->nx) * emf->ny * emf->nz;
    fwi_.n = static_cast<int>(nxyz_ * fwi_.npar);
    if (emf->verb) {
      printf("preco=%d (0=off; 1=Shin's preconditioner; 2=depth weighting; 3=triangle smoothing)\n", fwi_.preco);
      printf("Amplitude uncertainty: %g%%\n", emf->amp_perc * 100.);
      printf("noisefloorE=%g, noisefloorH=%g\n", emf->noisefloorE, emf->noisefloorH);
      printf("Tikhonov regularization: gamma1=%g\n", fwi_.gamma1);
      printf("TV regularization:       gamma2=%g\n", fwi_.gamma2);
      printf("----------- fwi init ------------\n");
    }
    // bathymetry (src/fg_fwi.c:86-99)
    const size_t nx = emf->nx, ny = emf->ny, nz = emf->nz, nf = emf->nfreq;
    emf->bathy = alloc2<float>(nx, ny);
    FILE* fp = fopen(fbathy, "rb");
    if (!fp) emmi_cli_err("cannot open file fbathy=%s\n", fbathy);
    const size_t got = fread(&emf->bathy[0][0], sizeof(float), nx * ny, fp);
    fclose(fp);
    if (got != nx * ny) emmi_cli_err("size parameter does not match the file-bathy!");
    emf->depthmax = emf->bathy[0][0];
    for (size_t i = 0; i < nx * ny; ++i) emf->depthmax = std::max(emf->bathy[0][i], emf->depthmax);
    if (emf->verb) printf("maximum water depth: depthmax=%g\n", emf->depthmax);
    // every rank: acquisition, observed data with noise, obs_<shot>.txt;
    // the noise generator starts fresh per rank as in a new process
    ranks_.resize(nproc);
    float mse = 0;
    for (iproc = 0; iproc < nproc; ++iproc) {
      set_verb();
      Rank& r = ranks_[iproc];
      acq_init(&r.acq, emf);
      if (emf->depthmax < r.acq.src_x3[0]) emmi_cli_err("coordinate error: source placed below seabed!");
      r.dobs = alloc3<emmi_c64>(r.acq.nrec, nf, emf->nchrec);
      r.dcal = alloc3<emmi_c64>(r.acq.nrec, nf, emf->nchrec);
      r.dres = alloc3<emmi_c64>(r.acq.nrec, nf, emf->nchrec);
      r.delta = alloc3<float>(r.acq.nrec, nf, emf->nchrec);
      r.rmse = alloc3<float>(r.acq.nrec, nf, emf->nchrec);
      bind(r);
      read_data(&r.acq, emf);
      emmi_rannor_reset();
      mse = mse + cal_uncertainty_noise(&r.acq, emf);
      char fname[32];
      snprintf(fname, sizeof fname, "obs_%04d.txt", r.acq.shot_idx[iproc]);
      write_data(&r.acq, emf, fname, emf->dobs_fd);
    }
    iproc = 0;
    set_verb();
    fwi_.mse = mse;
    if (emf->verb) {
      const int n = nproc * emf->nchrec * emf->nfreq * ranks_[0].acq.nrec;
      printf("addnoise=%d, target rmse=%e\n", emf->addnoise, std::sqrt(mse / n));
    }
    xref_.resize(fwi_.n);
    fwi_.xref = xref_.data();
    fwi_.grad = alloc4<float>(nx, ny, nz, fwi_.npar);
    if (fwi_.preco == 1) fwi_.phess = alloc4<float>(nx, ny, nz, fwi_.npar);
    fwi_.iter = 0;
    fwi_.firstgrad = 1;
    fwi_.alpha = 1;
    // shots solved together per survey solve (batch_shots=, extension):
    // every right-hand side of a group is one slot of the solver (at most
    // 64), and the default keeps a group near 24 slots
    batch_shots_ = par_int("batch_shots", std::max(1, 24 / std::max(1, (int)nf)));
    if (batch_shots_ < 1 || batch_shots_ * (int)nf > 64)
      emmi_cli_err("batch_shots=%d: need 1 <= batch_shots * nfreq <= 64", batch_shots_);
    batch_shots_ = std::min(batch_shots_, nproc);
    // device buffers of the gradient evaluations, per shot of a group
    const size_t ebytes = sizeof(emmi_c64) * 3 * nxyz_;
    ef_.resize(nf * batch_shots_);
    ea_.resize(nf * batch_shots_);
    for (size_t f = 0; f < ef_.size(); ++f) {
      ef_[f] = emmi_device_alloc(ebytes);
      ea_[f] = emmi_device_alloc(ebytes);
      if (!ef_[f] || !ea_[f]) emmi_cli_err("device allocation of the model-grid fields failed");
    }
    grad_d_ = static_cast<float*>(emmi_device_alloc(sizeof(float) * fwi_.n));
    phess_d_ = fwi_.preco == 1 ? static_cast<float*>(emmi_device_alloc(sizeof(float) * fwi_.n)) : nullptr;
    if (!grad_d_ || (fwi_.preco == 1 && !phess_d_)) emmi_cli_err("device allocation of the gradient failed");
  }

  ~Inversion() {
    for (void* q : ef_) emmi_device_free(q);
    for (void* q : ea_) emmi_device_free(q);
    emmi_device_free(grad_d_);
    if (phess_d_) emmi_device_free(phess_d_);
    for (Rank& r : ranks_) {
      free3(r.dobs);
      free3(r.dcal);
      free3(r.dres);
      free3(r.delta);
      free3(r.rmse);
      acq_free(&r.acq);
    }
    emf_->dobs_fd = emf_->dcal_fd = emf_->dres_fd = nullptr;
    emf_->delta_emf = emf_->rmse = nullptr;
    free4(fwi_.grad);
    free4(fwi_.phess);
    free2(emf_->bathy);
    emf_->bathy = nullptr;
  }

  // src/do_fwi.c:43-240 for mode 1 (l-BFGS) and mode 2 (one gradient)
  void run() {
    emf_t* emf = emf_;
    LineSearch ls;
    const int niter = par_int("niter", 30);
    ls.nls = par_int("nls", 5);
    float tol = 1e-6f;
    getparfloat(const_cast<char*>("tol"), &tol);
    const int npair = par_int("npair", 5);
    ls.c1 = par_float("c1", 1e-4f);
    ls.c2 = par_float("c2", 0.9f);
    const float alpha0 = par_float("alpha", 1.);
    const int bound = par_int("bound", 0);
    const int method = par_int("method", 0);
    ls.verb = emf->verb;
    const int n = fwi_.n;
    std::vector<float> minpar, maxpar;
    if (bound) {
      const int j1 = countparval(const_cast<char*>("minpar")), j2 = countparval(const_cast<char*>("maxpar"));
      if (!j1) emmi_cli_err("Need lower bound of parameters minpar= vector");
      if (j1 != fwi_.npar) emmi_cli_err("npar=%d, must have length[idxpar]=npar", fwi_.npar);
      if (!j2) emmi_cli_err("Need upper bound of parameters maxpar= vector");
      if (j2 != fwi_.npar) emmi_cli_err("npar=%d, must have length[idxpar]=npar", fwi_.npar);
      minpar.resize(fwi_.npar);
      maxpar.resize(fwi_.npar);
      getparfloat(const_cast<char*>("minpar"), minpar.data());
      getparfloat(const_cast<char*>("maxpar"), maxpar.data());
    }
    if (emf->verb) {
      printf("---------------------------------------\n");
      printf("npar=%d\n", fwi_.npar);
      printf("**** method=%d (0=l-BFGS; 1=Gauss-Newton)\n", method);
      printf("**** niter=%d\n**** nls=%d\n**** bound=%d\n**** preco=%d\n", niter, ls.nls, bound, fwi_.preco);
    }
    fwi_.niter = niter;
    // the unknowns: log resistivities (src/do_fwi.c:110-121)
    std::vector<float> x(n), g(n), d(n), xmin, xmax;
    for (int j = 0; j < fwi_.npar; ++j)
      for (size_t i = 0; i < nxyz_; ++i) {
        if (idxpar_[j] == 1) x[i + nxyz_ * j] = static_cast<float>(std::log(static_cast<double>(emf->rho11[0][0][i])));
        if (idxpar_[j] == 2) x[i + nxyz_ * j] = static_cast<float>(std::log(static_cast<double>(emf->rho33[0][0][i])));
      }
    xref_ = x;
    fwi_.xref = xref_.data();
    if (bound) {
      // bounds below the seabed; the water and air keep their values
      xmin.resize(n);
      xmax.resize(n);
      for (int j = 0; j < fwi_.npar; ++j)
        for (int i3 = 0; i3 < emf->nz; ++i3)
          for (int i2 = 0; i2 < emf->ny; ++i2)
            for (int i1 = 0; i1 < emf->nx; ++i1) {
              const size_t i = i1 + emf->nx * (i2 + emf->ny * (i3 + static_cast<size_t>(emf->nz) * j));
              if (emf->oz + (i3 + 0.5) * emf->dz > emf->bathy[i2][i1] + emf->dz) {
                xmin[i] = static_cast<float>(std::log(static_cast<double>(minpar[j])));
                xmax[i] = static_cast<float>(std::log(static_cast<double>(maxpar[j])));
              } else {
                xmin[i] = xmax[i] = x[i];
              }
            }
    }
    const auto fg = [this](float* xx, float* gg) { return evaluate(xx, gg); };
    float fcost = evaluate(x.data(), g.data());
    if (emf->mode != 1) return;

    // l-BFGS with the bisection line search (src/do_fwi.c:151-225)
    (void)tol;  // the reference reads tol= for the loop but stops only at niter
    (void)alpha0;
    Lbfgs lb(n, npair);
    const float f0 = fcost;
    ls.fk = fcost;
    ls.igrad = 0;
    ls.ils = 0;
    ls.alpha = alpha0;
    if (emf->verb) {
      FILE* fp = fopen("iterate.txt", "w");
      if (!fp) emmi_cli_err("cannot open iterate.txt");
      fprintf(fp, "==========================================================\n");
      fprintf(fp, "l-BFGS memory length: %d\n", npair);
      fprintf(fp, "Maximum number of iterations: %d\n", niter);
      fprintf(fp, "Convergence tolerance: %3.2e\n", tol);
      fprintf(fp, "maximum number of line search: %d\n", ls.nls);
      fprintf(fp, "initial step length: alpha=%g\n", alpha0);
      fprintf(fp, "==========================================================\n");
      fprintf(fp, "iter    fk       fk/f0      ||gk||    alpha     nls   ngrad\n");
      fclose(fp);
    }
    int iter = 0;
    for (; iter < niter; ++iter) {
      fwi_.iter = iter;
      if (emf->verb) {
        printf("==========================================================\n");
        printf("# iter=%d  fk/f0=%g\n", iter, ls.fk / f0);
        const float gk = l2norm(n, g.data());
        FILE* fp = fopen("iterate.txt", "a");
        fprintf(fp, "%3d   %3.2e  %3.2e   %3.2e  %3.2e  %3d  %4d\n", iter, ls.fk, ls.fk / f0, gk, ls.alpha, ls.ils,
                ls.igrad);
        fclose(fp);
        fp = fopen("rmse_misfit.txt", iter == 0 ? "w" : "a");
        fprintf(fp, "%d \t %g\n", fwi_.iter, fwi_.mse);
        fclose(fp);
      }
      std::vector<float>& q = lb.q;
      if (iter == 0) {
        memcpy(q.data(), g.data(), sizeof(float) * n);
        if (fwi_.preco) precondition(emf, &fwi_, q.data());
        for (int j = 0; j < n; ++j) d[j] = -q[j];
      } else {
        lb.update(x.data(), g.data());
        if (fwi_.preco) {
          lb.first_loop(g.data());
          precondition(emf, &fwi_, q.data());
          // without a usable pair the reference keeps the previous direction
          if (lb.loop1) lb.second_loop(d.data());
        } else {
          lb.direction(g.data(), d.data());
        }
      }
      lb.save(x.data(), g.data());
      line_search(ls, n, x.data(), g.data(), d.data(), fg, bound ? xmin.data() : nullptr,
                  bound ? xmax.data() : nullptr);
      if (emf->verb) {
        char name[32];
        snprintf(name, sizeof name, "Rv_iteration_%02d", iter);
        write_bin(name, &emf->rho33[0][0][0], sizeof(float) * nxyz_);
        snprintf(name, sizeof name, "Rh_iteration_%02d", iter);
        write_bin(name, &emf->rho11[0][0][0], sizeof(float) * nxyz_);
      }
    }
    if (emf->verb && iter == niter) {
      FILE* fp = fopen("iterate.txt", "a");
      fprintf(fp, "==>Maximum iteration number reached!\n");
      fclose(fp);
    }
  }

 private:
  emf_t* emf_;
  emmi_params_t p_;
  int verb_set_, verb_par_;
  fwi_t fwi_{};
  std::vector<int> idxpar_;
  std::vector<float> xref_;
  size_t nxyz_ = 0;
  std::vector<Rank> ranks_;
  std::vector<void*> ef_, ea_;   // [shot of a group][frequency]
  int batch_shots_ = 1;
  float *grad_d_ = nullptr, *phess_d_ = nullptr;

  void set_verb() { emf_->verb = verb_set_ ? verb_par_ : (iproc == 0); }

  void bind(Rank& r) {
    emf_->dobs_fd = r.dobs;
    emf_->dcal_fd = r.dcal;
    emf_->dres_fd = r.dres;
    emf_->delta_emf = r.delta;
    emf_->rmse = r.rmse;
  }

  // one misfit and gradient evaluation at the log resistivities xx, all
  // ranks (src/fg_fwi.c:139-300); returns the scaled misfit, g its gradient
  float evaluate(float* xx, float* g) {
    emf_t* emf = emf_;
    const size_t nf = emf->nfreq;
    const int npar = fwi_.npar;
    emf->rhomin = emf->rhomax = static_cast<float>(std::exp(static_cast<double>(xx[0])));
    for (int j = 0; j < npar; ++j)
      for (size_t i = 0; i < nxyz_; ++i) {
        const float s = static_cast<float>(std::exp(static_cast<double>(xx[i + nxyz_ * j])));
        if (idxpar_[j] == 1) emf->rho11[0][0][i] = emf->rho22[0][0][i] = s;
        if (idxpar_[j] == 2) emf->rho33[0][0][i] = s;
        emf->rhomin = std::min(emf->rhomin, s);
        emf->rhomax = std::max(emf->rhomax, s);
      }
    if (emf->verb) printf("fg_fwi iter=%d: [rhomin,rhomax]=[%g,%g]\n", fwi_.iter, emf->rhomin, emf->rhomax);
    std::vector<float> gsum(fwi_.n, 0.0f), hsum(fwi_.preco == 1 ? fwi_.n : 0, 0.0f), part(fwi_.n);
    float fcost_dat = 0;
    int ndata0 = 0;  // rank 0's data count, times nproc below, as rank 0 forms it
    // the ranks in groups of batch_shots, each group one survey solve; every
    // step that forms a sum over ranks runs in rank order
    for (int r0 = 0; r0 < nproc; r0 += batch_shots_) {
      const int ns = std::min(batch_shots_, nproc - r0);
      std::vector<emf_t> emfs;
      std::vector<acq_t*> acqs;
      for (int j = 0; j < ns; ++j) {
        iproc = r0 + j;
        set_verb();
        bind(ranks_[iproc]);
        emfs.push_back(*emf);
        acqs.push_back(&ranks_[iproc].acq);
      }
      iproc = r0;
      set_verb();
      if (emf->verb) printf("--------step 1: forward modelling------------\n");
      // the forward and adjoint solves share grids, media and frequencies
      ShotGroup S(emfs, acqs, p_);
      std::vector<std::vector<void*>> ef(ns), ea(ns);
      for (int j = 0; j < ns; ++j) {
        ef[j].assign(ef_.begin() + (size_t)j * nf, ef_.begin() + (size_t)(j + 1) * nf);
        ea[j].assign(ea_.begin() + (size_t)j * nf, ea_.begin() + (size_t)(j + 1) * nf);
        for (size_t f = 0; f < nf; ++f)
          check(emmi_inject_source_dev(S.acq[j], &S.E[j], f, S.bs(j, f), nullptr), "inject_source");
      }
      S.solve("forward");
      for (int j = 0; j < ns; ++j) {
        for (size_t f = 0; f < nf; ++f) {
          check(emmi_extract_emf_data_dev(S.acq[j], &S.E[j], f, S.xs(j, f), S.bs(j, f), nullptr), "extract_emf_data");
          check(emmi_extract_electric_field_dev(S.acq[j], &S.E[j], f, S.xs(j, f), ef[j][f], nullptr),
                "extract_electric_field");
        }
        iproc = r0 + j;
        set_verb();
        acq_t* acq = S.acq[j];
        emf_t* e = &S.E[j];
        e->verb = emf->verb;
        char fname[32];
        snprintf(fname, sizeof fname, "syn_%04d.txt", acq->shot_idx[iproc]);
        write_data(acq, e, fname, e->dcal_fd);

        if (e->verb) printf("------step 2: compute fcost + adjoint source---------\n");
        const float fcost = emmi_data_misfit(acq, e);
        write_rmse(acq, e);
        const int n = e->nchrec * e->nfreq * acq->nrec;
        printf("isrc=%d, fcost=%g, rmse=%g\n", acq->shot_idx[iproc], fcost, std::sqrt(2. * fcost / n));
        fcost_dat = fcost_dat + fcost;
        if (iproc == 0) ndata0 = n;
        for (size_t f = 0; f < nf; ++f)
          check(emmi_inject_adj_source_dev(acq, e, f, S.xs(j, f), S.bs(j, f), nullptr), "inject_adj_source");
      }
      iproc = r0;
      set_verb();
      if (emf->verb) printf("-----------step3: adjoint modelling------------\n");
      S.solve("adjoint");
      for (int j = 0; j < ns; ++j) {
        iproc = r0 + j;
        set_verb();
        for (size_t f = 0; f < nf; ++f)
          check(emmi_extract_electric_field_dev(S.acq[j], &S.E[j], f, S.xs(j, f), ea[j][f], nullptr),
                "extract_electric_field");
        if (emf->verb) printf("----------step 4: compute gradient --------------\n");
        check(emmi_build_gradient_dev(S.acq[j], &S.E[j], &fwi_, ef[j].data(), ea[j].data(), grad_d_, phess_d_,
                                      nullptr),
              "build_gradient");
        // the sums over ranks (MPI_Allreduce in the reference), rank order
        if (emmi_memcpy_d2h(part.data(), grad_d_, sizeof(float) * fwi_.n)) emmi_cli_err("gradient copy failed");
        for (int i = 0; i < fwi_.n; ++i) gsum[i] = gsum[i] + part[i];
        if (fwi_.preco == 1) {
          if (emmi_memcpy_d2h(part.data(), phess_d_, sizeof(float) * fwi_.n)) emmi_cli_err("pseudo-Hessian copy failed");
          for (int i = 0; i < fwi_.n; ++i) hsum[i] = hsum[i] + part[i];
        }
      }
    }
    iproc = 0;
    set_verb();
    if (fwi_.preco == 1) memcpy(&fwi_.phess[0][0][0][0], hsum.data(), sizeof(float) * fwi_.n);
    memcpy(g, gsum.data(), sizeof(float) * fwi_.n);
    fwi_.fcost_dat = fcost_dat;
    fwi_.mse = static_cast<float>(std::sqrt(2. * fcost_dat / (ndata0 * nproc)));
    if (emf->verb) printf("fcost_dat=%g, rmse=%g\n", fwi_.fcost_dat, fwi_.mse);
    if (emf->verb) {
      // the data gradient before regularisation (src/fg_fwi.c:276-285)
      write_bin("gradient_dat_Rv", g, sizeof(float) * nxyz_);
      if (npar > 1) write_bin("gradient_dat_Rh", g + nxyz_, sizeof(float) * nxyz_);
    }
    fg_mod_reg(&ranks_[0].acq, emf, &fwi_, xx, g);
    fwi_.fcost = fwi_.fcost_dat + fwi_.fcost_mod;
    if (emf->verb) {
      printf("fcost_dat=%g\n", fwi_.fcost_dat);
      printf("fcost_mod=%g\n", fwi_.fcost_mod);
      printf("fcost=%g\n", fwi_.fcost);
    }
    // l-BFGS is not scale invariant: the first gradient of an inversion fixes
    // a scaling of misfit and gradient (src/fg_fwi.c:294-305)
    if (emf->mode == 1 && fwi_.firstgrad) {
      float s1 = 0, s2 = 0;
      for (int k = 0; k < fwi_.n; ++k) {  // float sums of double moduli
        s1 = static_cast<float>(s1 + std::fabs(static_cast<double>(xx[k])));
        s2 = static_cast<float>(s2 + std::fabs(static_cast<double>(g[k])));
      }
      fwi_.alpha = static_cast<float>(2e-3 * s1 / s2);
      if (emf->verb) printf("|x|_1=%g, |g|_1=%g, scaling=%g\n", s1, s2, fwi_.alpha);
    }
    for (int k = 0; k < fwi_.n; ++k) g[k] *= fwi_.alpha;
    if (fwi_.firstgrad) fwi_.firstgrad = 0;
    return fwi_.fcost * fwi_.alpha;
  }
};
}  // namespace

int main(int argc, char** argv) {
  setvbuf(stdout, nullptr, _IOLBF, 0);
  setvbuf(stderr, nullptr, _IOLBF, 0);
  initargs(argc, argv);
  emf_t emf{};
  emf.mode = par_int("mode", 0);
  int verb_par = 1;
  const int verb_set = getparint(const_cast<char*>("verb"), &verb_par);
  emf.verb = verb_set ? verb_par : 1;
  // the ranks of the job: one per entry of shots=, else nproc= (default 1)
  nproc = countparval(const_cast<char*>("shots"));
  if (nproc == 0) nproc = par_int("nproc", 1);
  if (nproc < 1) emmi_cli_err("nproc=%d must be at least 1", nproc);
  iproc = 0;
  if (emf.verb) {
    char now[128];
    const time_t t = time(nullptr);
    strftime(now, sizeof now, "%d-%b-%Y %H:%M:%S", localtime(&t));
    printf("  Current date and time: %s\n", now);
    printf("=====================================================\n");
    printf("    EM modelling and inversion, libEMMI_MGFD on MI355X\n");
    printf("=====================================================\n");
    printf(emf.mode == 0 ? "Task: Electromagnetic forward modeling \n"
                         : emf.mode == 1 ? "Task: Inversion for medium parameters \n"
                                         : "Task: Compute CSEM inversion gradient\n");
    printf("=====================================================\n");
  }
  if (emf.mode != 0 && par_int("method", 0) != 0)
    emmi_cli_err("method=%d: only l-BFGS (method=0) is provided; the Gauss-Newton option needs a Hessian-vector "
                 "product that the reference does not implement either (src/fg_fwi.c:307-311)",
                 par_int("method", 0));
  emf_init(&emf);
  const emmi_params_t p = solver_params(emf.verb);
  if (emf.mode == 0) {
    for (iproc = 0; iproc < nproc; ++iproc) {
      emf.verb = verb_set ? verb_par : (iproc == 0);
      acq_t acq{};
      acq_init(&acq, &emf);
      do_modelling(&acq, &emf, p);
      acq_free(&acq);
    }
    iproc = 0;
  } else {
    Inversion inv(&emf, p, verb_set, emf.verb);
    inv.run();
  }
  emf_free(&emf);
  return 0;
}
