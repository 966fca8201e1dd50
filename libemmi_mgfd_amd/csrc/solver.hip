// solver.hip -- host orchestration of the multigrid Maxwell solve.
//
// Control flow follows src/gmg.c:1586-1668 (V, F, W cycles) and
// src/gmg.c:1894-1932 (gmg_apply); the coarse hierarchies of the three
// semi-coarsening patterns (src/gmg.c:1752-1756: axis icycle%3 kept) are built
// once per solver instead of every cycle.  Smoothing is line block
// Gauss-Seidel along each axis coarsened on the next level
// (src/gmg.c:1301-1313) with the lines swept in four colours; on the coarsest
// level the lines run along the axis that is never coarsened, which makes the
// coarsest solve exact when a single line remains.
#include "solver.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <stdexcept>
#include <string>

namespace emmi {

void check_hip(hipError_t e, const char* what) {
  if (e != hipSuccess) {
    char buf[512];
    snprintf(buf, sizeof buf, "HIP error in %s: %s", what, hipGetErrorString(e));
    throw std::runtime_error(buf);
  }
}

#define HIPCHECK(x) check_hip((x), #x)

void HostAxis::from_nodes(const std::vector<double>& nodes) {
  const int n = (int)nodes.size() - 1;
  x = nodes;
  xs.assign(n + 1, 0.0);
  d.assign(n + 1, 0.0);
  ds.assign(n + 1, 0.0);
  for (int i = 0; i < n; i++) {
    ds[i] = x[i + 1] - x[i];
    xs[i] = 0.5 * (x[i] + x[i + 1]);
  }
  xs[n] = x[n] + 0.5 * ds[n - 1];
  ds[n] = xs[n] - x[n];
  for (int i = 1; i <= n; i++) d[i] = xs[i] - xs[i - 1];
  d[0] = xs[0] - x[0];
}

template <typename T>
Level<T> DevLevel<T>::view(double omu) const {
  Level<T> L;
  for (int a = 0; a < 3; a++) {
    L.n[a] = n[a];
    L.W[a] = geo + (5 * a + 0) * gstride;
    L.IW[a] = geo + (5 * a + 1) * gstride;
    L.H[a] = geo + (5 * a + 2) * gstride;
    L.X[a] = geo + (5 * a + 3) * gstride;
  }
  L.es[0] = 1;
  L.es[1] = n[0] + 1;
  L.es[2] = (long)(n[0] + 1) * (n[1] + 1);
  L.nc = L.es[2] * (n[2] + 1);
  L.sbar = sbar;
  L.omu = omu;
  return L;
}

template <typename T>
Transfer<T> DevLevel<T>::transfer() const {
  Transfer<T> X;
  for (int a = 0; a < 3; a++) {
    X.sc[a] = sc[a];
    X.wr[a] = wr + a * wstride;
  }
  return X;
}

template struct DevLevel<double>;
template struct DevLevel<float>;

static int depth_of(int n) {
  // src/gmg.c:1688-1699
  int l = 0;
  while (n % 2 == 0 && n > 1) { l++; n >>= 1; }
  if (n > 2) l++;
  return l;
}

static void* dalloc(size_t bytes, std::vector<void*>& owned) {
  void* p = nullptr;
  HIPCHECK(hipMalloc(&p, bytes ? bytes : 16));
  owned.push_back(p);
  return p;
}

template <> Hier<double>& Solver::hier<double>() { return h64_; }
template <> Hier<float>& Solver::hier<float>() { return h32_; }

Solver::Solver(int n1, int n2, int n3, const float* x1, const float* x2, const float* x3, const float* sig11,
               const float* sig22, const float* sig33, int sig_on_device, float omega, const SolveParams& p,
               hipStream_t stream)
    : p_(p), s_(stream) {
  n_[0] = n1; n_[1] = n2; n_[2] = n3;
  for (int a = 0; a < 3; a++)
    if (n_[a] < 2) throw std::runtime_error("multigrid size must be at least 2 cells per axis");
  nc0_ = (long)(n1 + 1) * (n2 + 1) * (n3 + 1);
  lmax_ = std::max(depth_of(n1), std::max(depth_of(n2), depth_of(n3)));
  if (lmax_ < 1) lmax_ = 1;
  npat_ = p_.isemicoarsen ? 3 : 1;
  // I omega mu0 with omega stored as float (src/gmg.c:1709)
  omu_ = (double)omega * (4. * 3.14159265358979323846264 * 1e-7);

  // ---- host geometry of all levels and patterns ----
  const float* xf[3] = {x1, x2, x3};
  std::array<HostAxis, 3> a0;
  for (int a = 0; a < 3; a++) {
    std::vector<double> nodes(n_[a] + 1);
    for (int i = 0; i <= n_[a]; i++) nodes[i] = xf[a][i];
    a0[a].from_nodes(nodes);
  }
  dims_.assign(npat_, {});
  scs_.assign(npat_, {});
  axes_.assign(npat_, {});
  for (int c = 0; c < npat_; c++) {
    dims_[c].push_back({n1, n2, n3});
    scs_[c].push_back({1, 1, 1});
    axes_[c].push_back(a0);
    for (int lev = 1; lev < lmax_; lev++) {
      const auto& nf = dims_[c][lev - 1];
      std::array<int, 3> sc, n;
      for (int a = 0; a < 3; a++) sc[a] = (nf[a] % 2 == 0 && nf[a] > 2) ? 2 : 1;
      if (p_.isemicoarsen) sc[c] = 1;
      for (int a = 0; a < 3; a++) n[a] = nf[a] / sc[a];
      std::array<HostAxis, 3> ax;
      for (int a = 0; a < 3; a++) {
        std::vector<double> nodes(n[a] + 1);
        for (int i = 0; i <= n[a]; i++) nodes[i] = axes_[c][lev - 1][a].x[sc[a] * i];
        ax[a].from_nodes(nodes);
      }
      dims_[c].push_back(n);
      scs_[c].push_back(sc);
      axes_[c].push_back(ax);
    }
  }

  // ---- medium: sigma*vol cells on the device, all levels and patterns ----
  const long ncell0 = (long)n1 * n2 * n3;
  std::vector<void*> tmp;
  std::vector<std::vector<double*>> cells(npat_ * lmax_);  // [pat*lmax+lev][3]
  {
    double* W[3];
    for (int a = 0; a < 3; a++) {
      W[a] = (double*)dalloc(sizeof(double) * (n_[a] + 1), tmp);
      HIPCHECK(hipMemcpyAsync(W[a], a0[a].ds.data(), sizeof(double) * (n_[a] + 1), hipMemcpyHostToDevice, s_));
    }
    const float* sg[3] = {sig11, sig22, sig33};
    std::vector<double*> c0(3);
    for (int q = 0; q < 3; q++) {
      const float* src = sg[q];
      if (!sig_on_device) {
        float* d = (float*)dalloc(sizeof(float) * ncell0, tmp);
        HIPCHECK(hipMemcpyAsync(d, src, sizeof(float) * ncell0, hipMemcpyHostToDevice, s_));
        src = d;
      }
      c0[q] = (double*)dalloc(sizeof(double) * ncell0, tmp);
      launch_sigma_vol0(n1, n2, n3, W[0], W[1], W[2], src, c0[q], s_);
    }
    for (int c = 0; c < npat_; c++) {
      cells[c * lmax_ + 0] = c0;
      for (int lev = 1; lev < lmax_; lev++) {
        const auto& n = dims_[c][lev];
        const auto& nf = dims_[c][lev - 1];
        std::vector<double*> cl(3);
        for (int q = 0; q < 3; q++) {
          cl[q] = (double*)dalloc(sizeof(double) * (size_t)n[0] * n[1] * n[2], tmp);
          launch_coarsen_cells(nf[0], nf[1], n[0], n[1], n[2], scs_[c][lev].data(), cells[c * lmax_ + lev - 1][q], cl[q], s_);
        }
        cells[c * lmax_ + lev] = cl;
      }
    }
  }

  if (p_.precision == 0) {
    build_hier<double>(h64_, cells, lmax_);
  } else {
    build_hier<float>(h32_, cells, lmax_);
    // FP64 level-0 operator and state of the refinement
    Hier<double> one;
    build_hier<double>(one, cells, 1);
    l0_64_ = one.pat[0][0];
    u64_ = one.u[0];
    f64_ = one.f[0];
    r64_ = one.r[0];
  }
  if (p_.precision == 0) {
    u64_ = h64_.u[0];
    f64_ = h64_.f[0];
    r64_ = h64_.r[0];
    l0_64_ = h64_.pat[0][0];
  }

  // scratch of the line solves: max over levels, patterns, axes and colours
  size_t sreal = 0;
  for (int c = 0; c < npat_; c++)
    for (int lev = 0; lev < lmax_; lev++)
      for (int a = 0; a < 3; a++)
        for (int col = 0; col < 4; col++) sreal = std::max(sreal, line_scratch_doubles(dims_[c][lev].data(), a, col));
  lscr_ = (double*)dalloc(sizeof(double) * sreal, owned_);
  facs_.assign((size_t)npat_ * lmax_ * 3, LineFactors());

  partial_ = (double*)dalloc(sizeof(double) * residual_blocks(nc0_), owned_);
  dnorm_ = (double*)dalloc(sizeof(double) * 2, owned_);
  HIPCHECK(hipHostMalloc((void**)&hnorm_, sizeof(double) * 2, hipHostMallocDefault));
  HIPCHECK(hipStreamSynchronize(s_));
  for (void* t : tmp) HIPCHECK(hipFree(t));
}

Solver::~Solver() {
  hipStreamSynchronize(s_);
  for (void* p : owned_) hipFree(p);
  if (hnorm_) hipHostFree(hnorm_);
}

template <typename T>
void Solver::build_hier(Hier<T>& H, const std::vector<std::vector<double*>>& cells, int nlev) {
  H.pat.assign(npat_, std::vector<DevLevel<T>>(nlev));
  std::vector<long> maxnc(nlev, 0);
  for (int c = 0; c < npat_; c++)
    for (int lev = 0; lev < nlev; lev++) {
      DevLevel<T>& L = H.pat[c][lev];
      const auto& n = dims_[c][lev];
      for (int a = 0; a < 3; a++) { L.n[a] = n[a]; L.sc[a] = scs_[c][lev][a]; }
      L.nc = (long)(n[0] + 1) * (n[1] + 1) * (n[2] + 1);
      if (L.nc > maxnc[lev]) maxnc[lev] = L.nc;
      // level 0 is shared by the patterns
      if (c > 0 && lev == 0) { L = H.pat[0][0]; continue; }
      const int nmax = std::max(n[0], std::max(n[1], n[2]));
      L.gstride = nmax + 1;
      std::vector<T> g(15 * (size_t)L.gstride, T(0));
      for (int a = 0; a < 3; a++) {
        const HostAxis& ax = axes_[c][lev][a];
        for (int i = 0; i <= n[a]; i++) {
          g[(5 * a + 0) * L.gstride + i] = (T)ax.ds[i];
          g[(5 * a + 1) * L.gstride + i] = (T)(1.0 / ax.ds[i]);
          g[(5 * a + 2) * L.gstride + i] = i > 0 ? (T)(0.5 * (ax.ds[i] + ax.ds[i - 1])) : T(0);
          g[(5 * a + 3) * L.gstride + i] = (T)ax.x[i];
          g[(5 * a + 4) * L.gstride + i] = (T)ax.d[i];
        }
      }
      L.geo = (T*)dalloc(sizeof(T) * g.size(), owned_);
      HIPCHECK(hipMemcpyAsync(L.geo, g.data(), sizeof(T) * g.size(), hipMemcpyHostToDevice, s_));
      if (lev > 0) {
        // weights of the finer level's nodes towards this level (src/gmg.c:1478-1488)
        const auto& nf = dims_[c][lev - 1];
        const int nfmax = std::max(nf[0], std::max(nf[1], nf[2]));
        L.wstride = nfmax + 1;
        std::vector<T> w(3 * (size_t)L.wstride, T(0));
        for (int a = 0; a < 3; a++) {
          const HostAxis& fa = axes_[c][lev - 1][a];
          const HostAxis& ca = axes_[c][lev][a];
          for (int q = 0; q <= nf[a]; q++) {
            const int i = L.sc[a] == 2 ? q / 2 : q;
            if (i + 1 > n[a]) continue;
            w[a * L.wstride + q] = (T)((fa.x[q] - ca.x[i]) / (ca.x[i + 1] - ca.x[i]));
          }
        }
        L.wr = (T*)dalloc(sizeof(T) * w.size(), owned_);
        HIPCHECK(hipMemcpyAsync(L.wr, w.data(), sizeof(T) * w.size(), hipMemcpyHostToDevice, s_));
      }
      L.sbar = (T*)dalloc(sizeof(T) * 3 * L.nc, owned_);
      const auto& cl = cells[c * lmax_ + lev];
      launch_edge_sigma<T>(n[0], n[1], n[2], cl[0], cl[1], cl[2], L.sbar, s_);
    }
  H.u.assign(nlev, nullptr);
  H.f.assign(nlev, nullptr);
  H.r.assign(nlev, nullptr);
  for (int lev = 0; lev < nlev; lev++) {
    const size_t bytes = sizeof(typename cvec<T>::type) * 3 * (size_t)maxnc[lev];
    H.u[lev] = dalloc(bytes, owned_);
    H.f[lev] = dalloc(bytes, owned_);
    H.r[lev] = dalloc(bytes, owned_);
    launch_zero<T>(H.u[lev], 3 * maxnc[lev], s_);
    launch_zero<T>(H.f[lev], 3 * maxnc[lev], s_);
    launch_zero<T>(H.r[lev], 3 * maxnc[lev], s_);
  }
}

// ---------------------------------------------------------------------------
// cycles
// ---------------------------------------------------------------------------
template <typename T>
bool Solver::coarsens(int pat, int lev) {
  if (lev >= lmax_ - 1) return false;
  const auto& sc = scs_[pat][lev + 1];
  return sc[0] * sc[1] * sc[2] > 1;
}

// factor records of (pattern, level, axis), allocated on first use and
// rebuilt when the frequency changed since they were made
template <typename T>
LineFactors& Solver::factors(int pat, int lev, int axis) {
  const int pe = lev == 0 ? 0 : pat;
  LineFactors& F = facs_[((size_t)pe * lmax_ + lev) * 3 + axis];
  if (F.omu == omu_) return F;
  const DevLevel<T>& L = hier<T>().pat[pe][lev];
  LinePack<T> P;
  P.njobs = 4;
  for (int col = 0; col < 4; col++) {
    if (!F.fac[col]) F.fac[col] = (double*)dalloc(sizeof(double) * std::max<size_t>(1, line_factor_doubles(L.n, axis, col)), owned_);
    LineJob<T>& J = P.job[col];
    J.L = L.view(omu_);
    J.u = nullptr;
    J.f = nullptr;
    J.fac = F.fac[col];
    J.scr = nullptr;
    line_job_colour(J, axis, col);
  }
  launch_line_factor<T>(P, axis, s_);
  F.omu = omu_;
  return F;
}

template <typename T>
void Solver::line_sweep_on(int pat, int lev, int axis, int iter, void* u, const void* f) {
  const LineFactors& F = factors<T>(pat, lev, axis);
  const DevLevel<T>& L = hier<T>().pat[pat][lev];
  LinePack<T> P;
  P.njobs = 1;
  LineJob<T>& J = P.job[0];
  J.L = L.view(omu_);
  J.u = (typename cvec<T>::type*)u;
  J.f = (const typename cvec<T>::type*)f;
  J.scr = lscr_;
  for (int k = 0; k < 4; k++) {
    const int colour = (iter % 2 == 1) ? 3 - k : k;
    J.fac = F.fac[colour];
    line_job_colour(J, axis, colour);
    launch_line_solve<T>(P, axis, s_);
  }
}

template void Solver::line_sweep_on<double>(int, int, int, int, void*, const void*);
template void Solver::line_sweep_on<float>(int, int, int, int, void*, const void*);

template <typename T>
void Solver::line_sweep(int pat, int lev, int axis, int iter) {
  Hier<T>& H = hier<T>();
  line_sweep_on<T>(pat, lev, axis, iter, H.u[lev], H.f[lev]);
}

// src/gmg.c:1301-1313
template <typename T>
void Solver::smooth(int pat, int lev, int iter) {
  if (p_.isemicoarsen && lev < lmax_ - 1) {
    const auto& sc = scs_[pat][lev + 1];
    for (int a = 0; a < 3; a++)
      if (sc[a] == 2) line_sweep<T>(pat, lev, a, iter);
  } else if (p_.isemicoarsen) {
    for (int s = 0; s < p_.coarse_sweeps; s++) line_sweep<T>(pat, lev, pat, iter + s);
  } else {
    for (int a = 0; a < 3; a++) line_sweep<T>(pat, lev, a, iter);
  }
}

bool Solver::check(double rn) {
  if (st_) st_->history.push_back(rn);
  if (p_.verb) printf("icycle=%d rnorm=%e\n", icycle_, rn);
  if (!std::isfinite(rn)) {
    if (st_) st_->status = 2;
    return true;
  }
  if (st_) st_->rnorm = rn;
  return rn < rnorm0_ * p_.tol;
}

double Solver::norm_of(const DevLevel<double>& L, const void* u, const void* f, void* r) {
  launch_residual<double>(L.view(omu_), u, f, r, partial_, s_);
  launch_sum_partials(partial_, residual_blocks(L.nc), dnorm_, s_);
  HIPCHECK(hipMemcpyAsync(hnorm_, dnorm_, sizeof(double), hipMemcpyDeviceToHost, s_));
  HIPCHECK(hipStreamSynchronize(s_));
  return std::sqrt(hnorm_[0]);
}

// restriction of the level's residual (or right-hand side when the guess is
// zero) followed by the recursive coarse solve and the prolongation; returns
// true when the level-0 convergence test stopped the solve
template <typename T>
bool Solver::coarse_correction(int pat, int lev, bool norm_here, bool zero_guess, int kind) {
  Hier<T>& H = hier<T>();
  const DevLevel<T>& L = H.pat[pat][lev];
  const DevLevel<T>& C = H.pat[pat][lev + 1];
  const void* src;
  if (lev == 0 && norm_here) {
    // level-0 residual with its norm (only in the FP64 direct form)
    const double rn = norm_of(h64_.pat[pat][0], h64_.u[0], h64_.f[0], h64_.r[0]);
    if (check(rn)) { stopped_ = true; return true; }
    src = H.r[lev];
  } else if (zero_guess) {
    src = H.f[lev];
  } else {
    launch_residual<T>(L.view(omu_), H.u[lev], H.f[lev], H.r[lev], nullptr, s_);
    src = H.r[lev];
  }
  launch_restrict<T>(L.view(omu_), C.view(omu_), C.transfer(), src, H.f[lev + 1], s_);
  launch_zero<T>(H.u[lev + 1], 3 * C.nc, s_);
  if (kind == 2) f_cycle<T>(pat, lev + 1, false, true);
  else if (kind == 3) { w_cycle<T>(pat, lev + 1, false, true); w_cycle<T>(pat, lev + 1, false, false); }
  else v_cycle<T>(pat, lev + 1, false, true);
  launch_prolong<T>(L.view(omu_), C.view(omu_), C.transfer(), H.u[lev + 1], H.u[lev], s_);
  return false;
}

// src/gmg.c:1586-1612
template <typename T>
bool Solver::v_cycle(int pat, int lev, bool norm_here, bool zero_guess) {
  for (int i = 0; i < p_.v1; i++) { smooth<T>(pat, lev, i); zero_guess = false; }
  if (coarsens<T>(pat, lev)) {
    if (coarse_correction<T>(pat, lev, norm_here && p_.cycleopt == 1, zero_guess, 1)) return true;
  }
  for (int i = 0; i < p_.v2; i++) smooth<T>(pat, lev, i);
  return false;
}

// src/gmg.c:1615-1640
template <typename T>
bool Solver::f_cycle(int pat, int lev, bool norm_here, bool zero_guess) {
  Hier<T>& H = hier<T>();
  if (lev == lmax_ - 1) {
    launch_zero<T>(H.u[lev], 3 * H.pat[pat][lev].nc, s_);
    zero_guess = true;
  } else if (coarsens<T>(pat, lev)) {
    if (coarse_correction<T>(pat, lev, norm_here && p_.cycleopt == 2, zero_guess, 2)) return true;
    zero_guess = false;
  }
  return v_cycle<T>(pat, lev, false, zero_guess);
}

// src/gmg.c:1643-1668
template <typename T>
bool Solver::w_cycle(int pat, int lev, bool norm_here, bool zero_guess) {
  for (int i = 0; i < p_.v1; i++) { smooth<T>(pat, lev, i); zero_guess = false; }
  if (coarsens<T>(pat, lev)) {
    if (coarse_correction<T>(pat, lev, norm_here && p_.cycleopt == 3, zero_guess, 3)) return true;
  }
  for (int i = 0; i < p_.v2; i++) smooth<T>(pat, lev, i);
  return false;
}

double Solver::vec_norm(const void* v) {
  // ||v|| over all entries: the residual of u = 0 is f itself
  launch_zero<double>(r64_, 3 * nc0_, s_);
  return norm_of(l0_64_, r64_, v, r64_);
}

int Solver::solve(const void* b, void* x, void* h, const void* x0, SolveStats* st) {
  SolveStats local;
  st_ = st ? st : &local;
  *st_ = SolveStats();
  stopped_ = false;
  const Level<double> V0 = l0_64_.view(omu_);
  const long nv = 3 * nc0_;
  // f = b * dual volume (src/gmg.c:1899-1910)
  launch_scale_rhs<double>(V0, l0_64_.D(0), l0_64_.D(1), l0_64_.D(2), b, f64_, s_);
  // ||f||: the reference's rnorm0 (its first residual, taken at u = 0)
  {
    void* z = u64_;
    launch_zero<double>(z, nv, s_);
    rnorm0_ = norm_of(l0_64_, z, f64_, r64_);
  }
  st_->rnorm0 = rnorm0_;
  if (x0) HIPCHECK(hipMemcpyAsync(u64_, x0, sizeof(double2) * nv, hipMemcpyDeviceToDevice, s_));

  const int kind = p_.cycleopt;
  for (icycle_ = 0; icycle_ < p_.ncycle; icycle_++) {
    const int pat = p_.isemicoarsen ? icycle_ % 3 : 0;
    if (p_.precision == 0) {
      // direct form: the level-0 residual and its norm open the cycle
      bool stop;
      if (kind == 1) stop = v_cycle<double>(pat, 0, true, false);
      else if (kind == 3) stop = w_cycle<double>(pat, 0, true, false);
      else stop = f_cycle<double>(pat, 0, true, false);
      if (stop || stopped_) break;
      if (lmax_ == 1 || !coarsens<double>(pat, 0)) {
        // no coarse level carries the norm test: test after the cycle
        const double rn = norm_of(l0_64_, u64_, f64_, r64_);
        if (check(rn)) break;
      }
    } else {
      // refinement: FP64 residual, FP32 cycle on it from zero, FP64 update
      const double rn = norm_of(l0_64_, u64_, f64_, r64_);
      if (check(rn)) break;
      launch_demote(r64_, h32_.f[0], nv, s_);
      launch_zero<float>(h32_.u[0], nv, s_);
      if (kind == 1) v_cycle<float>(pat, 0, false, true);
      else if (kind == 3) w_cycle<float>(pat, 0, false, true);
      else f_cycle<float>(pat, 0, false, true);
      launch_promote_add(u64_, h32_.u[0], nv, s_);
    }
  }
  st_->cycles = icycle_ < p_.ncycle ? icycle_ : p_.ncycle;
  if (st_->status == 0 && !(st_->rnorm < rnorm0_ * p_.tol)) st_->status = 1;
  if (st_->history.empty()) st_->status = 1;
  HIPCHECK(hipMemcpyAsync(x, u64_, sizeof(double2) * nv, hipMemcpyDeviceToDevice, s_));
  launch_curl_h<double>(V0, u64_, h, s_);
  HIPCHECK(hipStreamSynchronize(s_));
  HIPCHECK(hipGetLastError());
  st_ = nullptr;
  return 0;
}

}  // namespace emmi
