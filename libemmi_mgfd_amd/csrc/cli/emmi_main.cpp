// emmi_main.cpp -- the libEMMI_MGFD program (src/main.c:24-84) on the GPU
// solver: the same command lines, parameter files, input files and output
// files, with every solve of a shot run as one batch on the device.
//
//   mode=0  modelling (src/do_modelling.c:28-104): emf_<shot>.txt, and with
//           icheck=1 the fields Ex.bin .. Hz.bin
//   mode=2  one gradient of the data misfit (src/do_fwi.c:43-150 and
//           src/fg_fwi.c:43-137, 139-300 up to the data gradient):
//           obs_/syn_/rmse_<shot>.txt and gradient_dat_Rv, gradient_dat_Rh
//   mode=1  (the inversion loop) is not provided: the program says so and
//           exits with status 1.
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
// the same FP64 residual test), graphs=0|1, coarse_sweeps=N, nproc=N.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <ctime>
#include <vector>

#include "arrays.h"
#include "emmi_cli.h"

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
// mode 2: the data part of one misfit/gradient evaluation
// ---------------------------------------------------------------------------
struct Rank {
  acq_t acq{};
  emmi_c64 ***dobs = nullptr, ***dcal = nullptr, ***dres = nullptr;
  float ***delta = nullptr, ***rmse = nullptr;
};

// point the carrier at one rank's data arrays
void bind(emf_t* emf, Rank& r) {
  emf->dobs_fd = r.dobs;
  emf->dcal_fd = r.dcal;
  emf->dres_fd = r.dres;
  emf->delta_emf = r.delta;
  emf->rmse = r.rmse;
}

void do_gradient(emf_t* emf, const emmi_params_t& p, int verb_set, int verb_par) {
  // src/fg_fwi.c:55-76
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
  // src/do_fwi.c:65-72: the parameters of the gradient
  fwi_t fwi{};
  fwi.preco = par_int("preco", 1);
  fwi.npar = par_int("npar", 2);
  if (fwi.npar < 1 || fwi.npar > 8) emmi_cli_err("npar=%d outside 1..8", fwi.npar);
  std::vector<int> idxpar(fwi.npar);
  // the reference leaves idxpar unset when idxpar= is absent; here it
  // defaults to 1..npar (rho_h, then rho_v)
  for (int j = 0; j < fwi.npar; ++j) idxpar[j] = j + 1;
  if (countparval(const_cast<char*>("idxpar")) > fwi.npar) emmi_cli_err("idxpar= has more than npar=%d values", fwi.npar);
  getparint(const_cast<char*>("idxpar"), idxpar.data());
  fwi.idxpar = idxpar.data();
  const size_t nx = emf->nx, ny = emf->ny, nz = emf->nz, nxyz = nx * ny * nz, nf = emf->nfreq;
  fwi.n = static_cast<int>(nxyz * fwi.npar);

  // bathymetry (src/fg_fwi.c:86-99)
  emf->bathy = alloc2<float>(nx, ny);
  {
    FILE* fp = fopen(fbathy, "rb");
    if (!fp) emmi_cli_err("cannot open file fbathy=%s\n", fbathy);
    const size_t got = fread(&emf->bathy[0][0], sizeof(float), nx * ny, fp);
    fclose(fp);
    if (got != nx * ny) emmi_cli_err("size parameter does not match the file-bathy!");
  }
  emf->depthmax = emf->bathy[0][0];
  for (size_t i = 0; i < nx * ny; ++i) emf->depthmax = std::max(emf->bathy[0][i], emf->depthmax);
  if (verb_par) printf("maximum water depth: depthmax=%g\n", emf->depthmax);

  // every rank: its acquisition, observed data with noise, obs_<shot>.txt
  // (src/fg_fwi.c:100-116)
  std::vector<Rank> ranks(nproc);
  float mse = 0;
  for (iproc = 0; iproc < nproc; ++iproc) {
    emf->verb = verb_set ? verb_par : (iproc == 0);
    Rank& r = ranks[iproc];
    acq_init(&r.acq, emf);
    if (emf->depthmax < r.acq.src_x3[0]) emmi_cli_err("coordinate error: source placed below seabed!");
    r.dobs = alloc3<emmi_c64>(r.acq.nrec, nf, emf->nchrec);
    r.dcal = alloc3<emmi_c64>(r.acq.nrec, nf, emf->nchrec);
    r.dres = alloc3<emmi_c64>(r.acq.nrec, nf, emf->nchrec);
    r.delta = alloc3<float>(r.acq.nrec, nf, emf->nchrec);
    r.rmse = alloc3<float>(r.acq.nrec, nf, emf->nchrec);
    bind(emf, r);
    read_data(&r.acq, emf);
    emmi_rannor_reset();
    mse = mse + cal_uncertainty_noise(&r.acq, emf);
    char fname[32];
    snprintf(fname, sizeof fname, "obs_%04d.txt", r.acq.shot_idx[iproc]);
    write_data(&r.acq, emf, fname, emf->dobs_fd);
  }
  iproc = 0;
  emf->verb = verb_set ? verb_par : 1;
  if (emf->verb) {
    const int n = nproc * emf->nchrec * emf->nfreq * ranks[0].acq.nrec;
    printf("addnoise=%d, target rmse=%e\n", emf->addnoise, std::sqrt(mse / n));
  }

  // the model as the optimiser hands it over: log of the resistivities, and
  // back (src/do_fwi.c:110-121, src/fg_fwi.c:151-170)
  {
    std::vector<float> xx(nxyz * fwi.npar);
    for (int j = 0; j < fwi.npar; ++j)
      for (size_t i = 0; i < nxyz; ++i) {
        if (idxpar[j] == 1) xx[i + nxyz * j] = static_cast<float>(std::log(static_cast<double>(emf->rho11[0][0][i])));
        if (idxpar[j] == 2) xx[i + nxyz * j] = static_cast<float>(std::log(static_cast<double>(emf->rho33[0][0][i])));
      }
    emf->rhomin = emf->rhomax = static_cast<float>(std::exp(static_cast<double>(xx[0])));
    for (int j = 0; j < fwi.npar; ++j)
      for (size_t i = 0; i < nxyz; ++i) {
        const float s = static_cast<float>(std::exp(static_cast<double>(xx[i + nxyz * j])));
        if (idxpar[j] == 1) emf->rho11[0][0][i] = emf->rho22[0][0][i] = s;
        if (idxpar[j] == 2) emf->rho33[0][0][i] = s;
        emf->rhomin = std::min(emf->rhomin, s);
        emf->rhomax = std::max(emf->rhomax, s);
      }
    if (emf->verb) printf("fg_fwi iter=0: [rhomin,rhomax]=[%g,%g]\n", emf->rhomin, emf->rhomax);
  }

  // every rank: forward, misfit, adjoint, gradient (src/fg_fwi.c:179-268)
  std::vector<float> g(nxyz * fwi.npar, 0.0f), grad_h(nxyz * fwi.npar);
  const size_t ebytes = sizeof(emmi_c64) * 3 * nxyz;
  std::vector<void*> ef(nf), ea(nf);
  for (size_t f = 0; f < nf; ++f) {
    ef[f] = emmi_device_alloc(ebytes);
    ea[f] = emmi_device_alloc(ebytes);
    if (!ef[f] || !ea[f]) emmi_cli_err("device allocation of the model-grid fields failed");
  }
  float* grad_d = static_cast<float*>(emmi_device_alloc(sizeof(float) * nxyz * fwi.npar));
  float* phess_d = fwi.preco == 1 ? static_cast<float*>(emmi_device_alloc(sizeof(float) * nxyz * fwi.npar)) : nullptr;
  if (!grad_d || (fwi.preco == 1 && !phess_d)) emmi_cli_err("device allocation of the gradient failed");
  float fcost_dat = 0;
  int ndata = 0;
  for (iproc = 0; iproc < nproc; ++iproc) {
    emf->verb = verb_set ? verb_par : (iproc == 0);
    Rank& r = ranks[iproc];
    acq_t* acq = &r.acq;
    bind(emf, r);
    if (emf->verb) printf("--------step 1: forward modelling------------\n");
    regrid_init(acq, emf);
    {
      // the forward and adjoint solves share the grid, the medium and the
      // frequencies, so one solver serves both
      ShotSolve S(emf, p);
      for (int f = 0; f < emf->nfreq; ++f) check(emmi_inject_source_dev(acq, emf, f, S.b[f], nullptr), "inject_source");
      S.solve("forward");
      for (int f = 0; f < emf->nfreq; ++f) {
        check(emmi_extract_emf_data_dev(acq, emf, f, S.x[f], S.b[f], nullptr), "extract_emf_data");
        check(emmi_extract_electric_field_dev(acq, emf, f, S.x[f], ef[f], nullptr), "extract_electric_field");
      }
      char fname[32];
      snprintf(fname, sizeof fname, "syn_%04d.txt", acq->shot_idx[iproc]);
      write_data(acq, emf, fname, emf->dcal_fd);

      if (emf->verb) printf("------step 2: compute fcost + adjoint source---------\n");
      const float fcost = emmi_data_misfit(acq, emf);
      write_rmse(acq, emf);
      const int n = emf->nchrec * emf->nfreq * acq->nrec;
      printf("isrc=%d, fcost=%g, rmse=%g\n", acq->shot_idx[iproc], fcost, std::sqrt(2. * fcost / n));
      fcost_dat = fcost_dat + fcost;
      ndata += n;

      if (emf->verb) printf("-----------step3: adjoint modelling------------\n");
      for (int f = 0; f < emf->nfreq; ++f)
        check(emmi_inject_adj_source_dev(acq, emf, f, S.x[f], S.b[f], nullptr), "inject_adj_source");
      S.solve("adjoint");
      for (int f = 0; f < emf->nfreq; ++f)
        check(emmi_extract_electric_field_dev(acq, emf, f, S.x[f], ea[f], nullptr), "extract_electric_field");

      if (emf->verb) printf("----------step 4: compute gradient --------------\n");
      check(emmi_build_gradient_dev(acq, emf, &fwi, ef.data(), ea.data(), grad_d, phess_d, nullptr), "build_gradient");
      if (emmi_memcpy_d2h(grad_h.data(), grad_d, sizeof(float) * grad_h.size()))
        emmi_cli_err("copy of the gradient to the host failed");
      for (size_t i = 0; i < g.size(); ++i) g[i] = g[i] + grad_h[i];
    }
    regrid_free(emf);
  }
  iproc = 0;
  emf->verb = verb_set ? verb_par : 1;
  const float rms = static_cast<float>(std::sqrt(2. * fcost_dat / ndata));
  if (emf->verb) {
    printf("fcost_dat=%g, rmse=%g\n", fcost_dat, rms);
    // the data gradient before regularisation (src/fg_fwi.c:276-285)
    write_bin("gradient_dat_Rv", g.data(), sizeof(float) * nxyz);
    if (fwi.npar > 1) write_bin("gradient_dat_Rh", g.data() + nxyz, sizeof(float) * nxyz);
  }
  for (void* q : ef) emmi_device_free(q);
  for (void* q : ea) emmi_device_free(q);
  emmi_device_free(grad_d);
  if (phess_d) emmi_device_free(phess_d);
  for (Rank& r : ranks) {
    free3(r.dobs);
    free3(r.dcal);
    free3(r.dres);
    free3(r.delta);
    free3(r.rmse);
    acq_free(&r.acq);
  }
  emf->dobs_fd = emf->dcal_fd = emf->dres_fd = nullptr;
  emf->delta_emf = emf->rmse = nullptr;
  free2(emf->bathy);
  emf->bathy = nullptr;
}
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
  if (emf.mode == 1)
    emmi_cli_err("mode=1 (the inversion loop) is not provided by this program; mode=0 and mode=2 are");
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
    do_gradient(&emf, p, verb_set, emf.verb);
  }
  emf_free(&emf);
  return 0;
}
