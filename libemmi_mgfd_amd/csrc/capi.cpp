// capi.cpp -- C-ABI of the solver: the reference's gmg_init/gmg_apply/gmg_free
// and the extended emmi_* API.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>
#include <algorithm>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

#include "emmi_mgfd.h"
#include "solver.h"
#include "timing.h"

using namespace emmi;

static thread_local std::string g_err;

static void set_err(const char* m) { g_err = m ? m : ""; }

// getpar hooks: the reference's host program provides getparint and
// getpardouble (src/cstd.c); without one the defaults apply
extern "C" {
__attribute__((weak)) int getparint(char* name, int* p);
__attribute__((weak)) int getpardouble(char* name, double* p);
}

static int par_int(const char* name, int def) {
  int v = def;
  if (getparint && getparint(const_cast<char*>(name), &v)) return v;
  return def;
}
static double par_double(const char* name, double def) {
  double v = def;
  if (getpardouble && getpardouble(const_cast<char*>(name), &v)) return v;
  return def;
}

static SolveParams to_params(const emmi_params_t* p) {
  SolveParams q;
  if (p) {
    q.cycleopt = p->cycleopt;
    q.ncycle = p->ncycle;
    q.v1 = p->v1;
    q.v2 = p->v2;
    q.isemicoarsen = p->isemicoarsen;
    q.tol = p->tol;
    q.precision = p->precision;
    q.coarse_sweeps = p->coarse_sweeps;
    q.verb = p->verb;
    q.graphs = p->graphs;
  }
  return q;
}

static void validate(const SolveParams& q) {
  if (q.cycleopt < 1 || q.cycleopt > 3) throw std::runtime_error("cycleopt must be 1, 2 or 3");
  if (q.ncycle < 0 || q.v1 < 0 || q.v2 < 0) throw std::runtime_error("negative cycle or smoothing count");
  if (q.precision != 0 && q.precision != 1) throw std::runtime_error("precision must be 0 or 1");
  if (q.coarse_sweeps < 1) throw std::runtime_error("coarse_sweeps must be >= 1");
}

struct emmi_solver {
  Solver* s = nullptr;
  hipStream_t stream = nullptr;
  std::vector<std::vector<double>> hist;   // per slot, of the last solve
  int precision = 0;
};

extern "C" {

void emmi_default_params(emmi_params_t* p) {
  SolveParams q;
  p->cycleopt = q.cycleopt;
  p->ncycle = q.ncycle;
  p->v1 = q.v1;
  p->v2 = q.v2;
  p->isemicoarsen = q.isemicoarsen;
  p->tol = q.tol;
  p->precision = q.precision;
  p->coarse_sweeps = q.coarse_sweeps;
  p->verb = q.verb;
  p->graphs = q.graphs;
}

const char* emmi_last_error(void) { return g_err.c_str(); }

emmi_solver_t* emmi_solver_create_batch(int n1, int n2, int n3, const float* x1, const float* x2, const float* x3,
                                        const float* sig11, const float* sig22, const float* sig33, int sig_on_device,
                                        const float* omegas, int nslots, const emmi_params_t* p) {
  emmi_solver* h = new emmi_solver;
  try {
    SolveParams q = to_params(p);
    validate(q);
    if (n1 < 2 || n2 < 2 || n3 < 2) throw std::runtime_error("n1, n2, n3 must be >= 2");
    if (!x1 || !x2 || !x3 || !sig11 || !sig22 || !sig33) throw std::runtime_error("null grid or medium");
    if (!omegas || nslots < 1) throw std::runtime_error("need at least one frequency");
    if (nslots > Solver::MAX_SLOTS) throw std::runtime_error("at most 64 right-hand sides per solver");
    check_hip(hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking), "hipStreamCreate");
    h->precision = q.precision;
    h->s = new Solver(n1, n2, n3, x1, x2, x3, sig11, sig22, sig33, sig_on_device, omegas, nslots, q, h->stream);
    h->hist.assign(nslots, {});
    return h;
  } catch (const std::exception& e) {
    set_err(e.what());
    if (h->stream) hipStreamDestroy(h->stream);
    delete h;
    return nullptr;
  }
}

emmi_solver_t* emmi_solver_create_survey(int n1, int n2, int n3, int nmed, const float* const* x1,
                                         const float* const* x2, const float* const* x3, const float* const* sig11,
                                         const float* const* sig22, const float* const* sig33, int sig_on_device,
                                         const int* slot_med, const float* omegas, int nslots,
                                         const emmi_params_t* p) {
  emmi_solver* h = new emmi_solver;
  try {
    SolveParams q = to_params(p);
    validate(q);
    if (n1 < 2 || n2 < 2 || n3 < 2) throw std::runtime_error("n1, n2, n3 must be >= 2");
    if (nmed < 1 || !x1 || !x2 || !x3 || !sig11 || !sig22 || !sig33) throw std::runtime_error("null grid or medium");
    for (int m = 0; m < nmed; m++)
      if (!x1[m] || !x2[m] || !x3[m] || !sig11[m] || !sig22[m] || !sig33[m])
        throw std::runtime_error("null grid or medium in the survey");
    if (!omegas || !slot_med || nslots < 1) throw std::runtime_error("need at least one right-hand side");
    if (nslots > Solver::MAX_SLOTS) throw std::runtime_error("at most 64 right-hand sides per solver");
    check_hip(hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking), "hipStreamCreate");
    h->precision = q.precision;
    h->s = new Solver(n1, n2, n3, nmed, x1, x2, x3, sig11, sig22, sig33, sig_on_device, slot_med, omegas, nslots, q,
                      h->stream);
    h->hist.assign(nslots, {});
    return h;
  } catch (const std::exception& e) {
    set_err(e.what());
    if (h->stream) hipStreamDestroy(h->stream);
    delete h;
    return nullptr;
  }
}

emmi_solver_t* emmi_solver_create(int n1, int n2, int n3, const float* x1, const float* x2, const float* x3,
                                  const float* sig11, const float* sig22, const float* sig33, int sig_on_device,
                                  float omega, const emmi_params_t* p) {
  return emmi_solver_create_batch(n1, n2, n3, x1, x2, x3, sig11, sig22, sig33, sig_on_device, &omega, 1, p);
}

int emmi_solver_solve_batch(emmi_solver_t* h, int n, const void* const* b, void* const* x, void* const* hout,
                            const void* const* x0, emmi_stats_t* st) {
  if (!h || !h->s) { set_err("null solver"); return EMMI_ERROR; }
  try {
    if (n < 1 || n > h->s->nslots()) throw std::runtime_error("batch size must be 1..nslots");
    if (!b || !x || !hout) throw std::runtime_error("null vector table");
    for (int k = 0; k < n; k++)
      if (!b[k] || !x[k] || !hout[k]) throw std::runtime_error("null device vector in the batch");
    std::vector<SolveStats> ss(n);
    h->s->solve_batch(n, b, x, hout, x0, ss.data());
    int worst = 0;
    for (int k = 0; k < n; k++) {
      h->hist[k] = ss[k].history;
      if (st) {
        st[k].cycles = ss[k].cycles;
        st[k].status = ss[k].status;
        st[k].rnorm0 = ss[k].rnorm0;
        st[k].rnorm = ss[k].rnorm;
        st[k].x0_rejected = ss[k].x0_rejected;
        st[k].x0_rnorm = ss[k].x0_rnorm;
      }
      worst = std::max(worst, ss[k].status);
    }
    return worst;
  } catch (const std::exception& e) {
    set_err(e.what());
    return EMMI_ERROR;
  }
}

int emmi_solver_solve(emmi_solver_t* h, const void* b, void* x, void* hout, const void* x0, emmi_stats_t* st) {
  return emmi_solver_solve_batch(h, 1, &b, &x, &hout, &x0, st);
}

int emmi_solver_set_omega_slot(emmi_solver_t* h, int slot, float omega) {
  if (!h || !h->s) { set_err("null solver"); return EMMI_ERROR; }
  try {
    h->s->set_omega(slot, omega);
    return 0;
  } catch (const std::exception& e) {
    set_err(e.what());
    return EMMI_ERROR;
  }
}

int emmi_solver_set_omega(emmi_solver_t* h, float omega) { return emmi_solver_set_omega_slot(h, 0, omega); }

int emmi_solver_set_ncycle(emmi_solver_t* h, int ncycle) {
  if (!h || !h->s) { set_err("null solver"); return EMMI_ERROR; }
  if (ncycle < 0) { set_err("negative cycle count"); return EMMI_ERROR; }
  h->s->set_ncycle(ncycle);
  return 0;
}

int emmi_solver_autotune(emmi_solver_t* h, int reps) {
  if (!h || !h->s) { set_err("null solver"); return EMMI_ERROR; }
  try {
    return h->s->autotune(reps);
  } catch (const std::exception& e) {
    set_err(e.what());
    return EMMI_ERROR;
  }
}

int emmi_solver_blend_guess(emmi_solver_t* h, int n, const void* const* b, const void* const* x1, void* const* xg,
                            double* beta) {
  if (!h || !h->s) { set_err("null solver"); return EMMI_ERROR; }
  if (n < 0 || n > h->s->nslots()) { set_err("blend_guess: slot count out of range"); return EMMI_ERROR; }
  if (n > 0 && (!b || !x1 || !xg)) { set_err("blend_guess: null vector table"); return EMMI_ERROR; }
  for (int k = 0; k < n; k++)
    if (!b[k] || !x1[k] || !xg[k]) { set_err("blend_guess: null device vector"); return EMMI_ERROR; }
  try {
    h->s->blend_guess(n, b, x1, xg, beta);
    return 0;
  } catch (const std::exception& e) {
    set_err(e.what());
    return EMMI_ERROR;
  }
}

int emmi_line_tuning(int* rows, double* us, int maxn) { return emmi::line_choices(rows, us, maxn); }

void emmi_line_tuning_clear(void) { emmi::clear_line_choices(); }

int emmi_solver_set_first_cycle(emmi_solver_t* h, int first) {
  if (!h || !h->s) { set_err("null solver"); return EMMI_ERROR; }
  if (first < 0) { set_err("negative cycle index"); return EMMI_ERROR; }
  h->s->set_first_cycle(first % 3);
  return 0;
}

int emmi_solver_history_slot(emmi_solver_t* h, int slot, double* out, int maxn) {
  if (!h || slot < 0 || slot >= (int)h->hist.size()) return 0;
  const std::vector<double>& v = h->hist[slot];
  const int n = (int)v.size();
  for (int i = 0; i < n && i < maxn; i++) out[i] = v[i];
  return n;
}

int emmi_solver_history(emmi_solver_t* h, double* out, int maxn) { return emmi_solver_history_slot(h, 0, out, maxn); }

int emmi_solver_slots(emmi_solver_t* h) { return (h && h->s) ? h->s->nslots() : 0; }

void emmi_solver_destroy(emmi_solver_t* h) {
  if (!h) return;
  delete h->s;
  if (h->stream) hipStreamDestroy(h->stream);
  delete h;
}

int emmi_solver_levels(emmi_solver_t* h, int* npat) {
  if (!h) return 0;
  if (npat) *npat = h->s->npatterns();
  return h->s->lmax();
}

void emmi_solver_dims(emmi_solver_t* h, int pattern, int lev, int* n123) {
  const int* d = h->s->dims(pattern, lev);
  n123[0] = d[0]; n123[1] = d[1]; n123[2] = d[2];
}

void* emmi_device_alloc(size_t bytes) {
  void* p = nullptr;
  if (hipMalloc(&p, bytes ? bytes : 16) != hipSuccess) { set_err("hipMalloc failed"); return nullptr; }
  return p;
}
void emmi_device_free(void* p) { if (p) hipFree(p); }
int emmi_memcpy_h2d(void* dst, const void* src, size_t bytes) {
  return hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice) == hipSuccess ? 0 : -1;
}
int emmi_memcpy_d2h(void* dst, const void* src, size_t bytes) {
  return hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost) == hipSuccess ? 0 : -1;
}
int emmi_memcpy_d2d(void* dst, const void* src, size_t bytes) {
  // complete before returning, whatever stream later reads dst
  if (hipMemcpy(dst, src, bytes, hipMemcpyDeviceToDevice) != hipSuccess) return -1;
  return hipDeviceSynchronize() == hipSuccess ? 0 : -1;
}
int emmi_device_sync(void) { return hipDeviceSynchronize() == hipSuccess ? 0 : -1; }

int emmi_set_line_kernel(int mode) { return emmi::set_line_kernel(mode); }

int emmi_solver_kernel_rates(emmi_solver_t* h, int reps, double* ms, double* bytes, int n) {
  if (!h || !h->s) { set_err("null solver"); return EMMI_ERROR; }
  try {
    if (n < 6 || !ms || !bytes) throw std::runtime_error("kernel_rates needs room for 6 entries");
    return h->s->kernel_rates(reps, ms, bytes);
  } catch (const std::exception& e) {
    set_err(e.what());
    return EMMI_ERROR;
  }
}

void emmi_timing_enable(int on) { timing::enable(on != 0); }
int emmi_timing_read(double* ms, long* count, double* bytes, int n) { return timing::read(ms, count, bytes, n); }
int emmi_timing_read_layout(double* bytes, int n) { return timing::read_layout(bytes, n); }

}  // extern "C"

// ---------------------------------------------------------------------------
// single-kernel checks
// ---------------------------------------------------------------------------
namespace {

template <typename T>
struct DevVec {
  void* p = nullptr;
  long n = 0;
  explicit DevVec(long n_) : n(n_) { check_hip(hipMalloc(&p, sizeof(typename cvec<T>::type) * (n ? n : 1)), "hipMalloc"); }
  ~DevVec() { if (p) hipFree(p); }
  void up(const emmi_c128* h) {
    std::vector<typename cvec<T>::type> t(n);
    for (long i = 0; i < n; i++) { t[i].x = (T)h[i].re; t[i].y = (T)h[i].im; }
    check_hip(hipMemcpy(p, t.data(), sizeof(t[0]) * n, hipMemcpyHostToDevice), "h2d");
  }
  void down(emmi_c128* h) {
    std::vector<typename cvec<T>::type> t(n);
    check_hip(hipMemcpy(t.data(), p, sizeof(t[0]) * n, hipMemcpyDeviceToHost), "d2h");
    for (long i = 0; i < n; i++) { h[i].re = t[i].x; h[i].im = t[i].y; }
  }
};

template <typename T>
int check_op(Solver& S, const Hier<T>& H, int op, int pat, int lev, int axis, int iter, const emmi_c128* in1,
             const emmi_c128* in2, emmi_c128* out) {
  hipStream_t s = S.stream();
  const DevLevel<T>& L = H.pat[pat][lev];
  const Level<T> V = L.view(S.omu());
  const long nv = 3 * L.nc;
  if (op == 0 || op == 1 || op == 5 || op == 6) {
    DevVec<T> u(nv), f(nv), r(nv);
    u.up(in1);
    f.up(in2);
    if (op == 0) {
      launch_residual<T>(V, u.p, f.p, r.p, nullptr, s);
      check_hip(hipStreamSynchronize(s), "sync");
      r.down(out);
    } else if (op == 6) {
      void* uu[1] = {u.p};
      const void* ff[1] = {f.p};
      const double om[1] = {S.omu()};
      launch_point_gs_batch<T>(&V, 1, uu, ff, om, iter, s);
      check_hip(hipStreamSynchronize(s), "sync");
      u.down(out);
    } else {
      S.line_sweep_on<T>(pat, lev, axis, iter, u.p, f.p);
      check_hip(hipStreamSynchronize(s), "sync");
      u.down(out);
    }
    return 0;
  }
  if (op == 2 || op == 3) {
    if (lev + 1 >= (int)H.pat[pat].size()) throw std::runtime_error("no coarser level");
    const DevLevel<T>& C = H.pat[pat][lev + 1];
    const Level<T> VC = C.view(S.omu());
    DevVec<T> fine(nv), coarse(3 * C.nc);
    if (op == 2) {
      fine.up(in1);
      launch_restrict<T>(V, VC, C.transfer(), fine.p, coarse.p, nullptr, s);
      check_hip(hipStreamSynchronize(s), "sync");
      coarse.down(out);
    } else {
      coarse.up(in1);
      fine.up(in2);
      launch_prolong<T>(V, VC, C.transfer(), coarse.p, fine.p, s);
      check_hip(hipStreamSynchronize(s), "sync");
      fine.down(out);
    }
    return 0;
  }
  if (op == 4) {
    DevVec<T> u(nv), h(nv);
    u.up(in1);
    launch_curl_h<T>(V, u.p, h.p, s);
    check_hip(hipStreamSynchronize(s), "sync");
    h.down(out);
    return 0;
  }
  throw std::runtime_error("unknown check op");
}

}  // namespace

extern "C" int emmi_solver_check_op(emmi_solver_t* h, int op, int pattern, int lev, int axis, int iter,
                                    const emmi_c128* in1, const emmi_c128* in2, emmi_c128* out) {
  if (!h || !h->s) { set_err("null solver"); return EMMI_ERROR; }
  try {
    if (pattern < 0 || pattern >= h->s->npatterns() || lev < 0 || lev >= h->s->lmax()) throw std::runtime_error("bad level");
    if (axis < 0 || axis > 2) throw std::runtime_error("bad axis");
    if (h->precision == 0) return check_op<double>(*h->s, h->s->h64(), op, pattern, lev, axis, iter, in1, in2, out);
    return check_op<float>(*h->s, h->s->h32(), op, pattern, lev, axis, iter, in1, in2, out);
  } catch (const std::exception& e) {
    set_err(e.what());
    return EMMI_ERROR;
  }
}

// ---------------------------------------------------------------------------
// reference entry points gmg_init / gmg_apply / gmg_free (src/gmg.c:1704-1932)
// ---------------------------------------------------------------------------
namespace {
struct LegacyState {
  emf_t* emf = nullptr;
  float omega = 0;
  emmi_params_t p;
  bool autotune = false;
  bool init = false;
};
LegacyState g_legacy;

[[noreturn]] void fatal(const char* m) {
  fprintf(stderr, "emmi_mgfd: %s\n", m);
  exit(EXIT_FAILURE);
}
}  // namespace

extern "C" void gmg_init(emf_t* emf, int ifreq) {
  g_legacy.emf = emf;
  g_legacy.omega = emf->omegas[ifreq];
  emmi_default_params(&g_legacy.p);
  g_legacy.p.cycleopt = par_int("cycleopt", 2);
  g_legacy.p.ncycle = par_int("ncycle", 30);
  g_legacy.p.v1 = par_int("v1", 1);
  g_legacy.p.v2 = par_int("v2", 1);
  g_legacy.p.tol = par_double("tol", 1e-6);
  g_legacy.p.isemicoarsen = par_int("isemicoarsen", 1);
  g_legacy.p.precision = par_int("precision", 0);
  g_legacy.p.coarse_sweeps = par_int("coarse_sweeps", 2);
  g_legacy.p.verb = emf->verb;
  g_legacy.p.graphs = par_int("graphs", 1);
  g_legacy.autotune = par_int("autotune", 0) != 0;
  g_legacy.init = true;
}

extern "C" void gmg_apply(int n, emmi_c128* b, emmi_c128* x) {
  if (!g_legacy.init) fatal("gmg_apply called before gmg_init");
  emf_t* e = g_legacy.emf;
  const long nv = 3L * (e->n1 + 1) * (e->n2 + 1) * (e->n3 + 1);
  if (n != nv) fatal("gmg_apply: n != 3(n1+1)(n2+1)(n3+1)");
  emmi_solver_t* s = emmi_solver_create(e->n1, e->n2, e->n3, e->x1, e->x2, e->x3, &e->sigma11[0][0][0],
                                        &e->sigma22[0][0][0], &e->sigma33[0][0][0], 0, g_legacy.omega, &g_legacy.p);
  if (!s) fatal(emmi_last_error());
  if (g_legacy.autotune && emmi_solver_autotune(s, 5) < 0) fatal(emmi_last_error());
  void *db = emmi_device_alloc(sizeof(emmi_c128) * nv), *dx = emmi_device_alloc(sizeof(emmi_c128) * nv);
  if (!db || !dx) fatal("device allocation failed");
  emmi_memcpy_h2d(db, b, sizeof(emmi_c128) * nv);
  emmi_stats_t st;
  const int rc = emmi_solver_solve(s, db, dx, db, nullptr, &st);
  if (rc == EMMI_ERROR) fatal(emmi_last_error());
  emmi_memcpy_d2h(x, dx, sizeof(emmi_c128) * nv);
  emmi_memcpy_d2h(b, db, sizeof(emmi_c128) * nv);
  emmi_device_free(db);
  emmi_device_free(dx);
  emmi_solver_destroy(s);
}

extern "C" void gmg_free(void) { g_legacy.init = false; }
