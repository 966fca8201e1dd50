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
#include "timing.h"

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

template <> Hier<double>& Solver::hier<double>(int m) { return meds_[m].h64; }
template <> Hier<float>& Solver::hier<float>(int m) { return meds_[m].h32; }


static constexpr double MU0 = 4. * 3.14159265358979323846264 * 1e-7;

Solver::Solver(int n1, int n2, int n3, const float* x1, const float* x2, const float* x3, const float* sig11,
               const float* sig22, const float* sig33, int sig_on_device, const float* omegas, int nslots,
               const SolveParams& p, hipStream_t stream)
    : p_(p), s_(stream) {
  std::vector<int> med(nslots > 0 ? nslots : 1, 0);
  init(n1, n2, n3, 1, &x1, &x2, &x3, &sig11, &sig22, &sig33, sig_on_device, med.data(), omegas, nslots);
}

Solver::Solver(int n1, int n2, int n3, int nmed, const float* const* x1, const float* const* x2,
               const float* const* x3, const float* const* sig11, const float* const* sig22,
               const float* const* sig33, int sig_on_device, const int* slot_med, const float* omegas, int nslots,
               const SolveParams& p, hipStream_t stream)
    : p_(p), s_(stream) {
  init(n1, n2, n3, nmed, x1, x2, x3, sig11, sig22, sig33, sig_on_device, slot_med, omegas, nslots);
}

void Solver::init(int n1, int n2, int n3, int nmed, const float* const* x1, const float* const* x2,
                  const float* const* x3, const float* const* sig11, const float* const* sig22,
                  const float* const* sig33, int sig_on_device, const int* slot_med, const float* omegas,
                  int nslots) {
  n_[0] = n1; n_[1] = n2; n_[2] = n3;
  for (int a = 0; a < 3; a++)
    if (n_[a] < 2) throw std::runtime_error("multigrid size must be at least 2 cells per axis");
  if (nslots < 1) throw std::runtime_error("a solver needs at least one right-hand-side slot");
  if (nslots > MAX_SLOTS) throw std::runtime_error("at most 64 right-hand-side slots per solver");
  if (nmed < 1) throw std::runtime_error("a solver needs at least one grid and medium");
  for (int k = 0; k < nslots; k++)
    if (slot_med[k] < 0 || slot_med[k] >= nmed) throw std::runtime_error("slot medium index out of range");
  nc0_ = (long)(n1 + 1) * (n2 + 1) * (n3 + 1);
  lmax_ = std::max(depth_of(n1), std::max(depth_of(n2), depth_of(n3)));
  if (lmax_ < 1) lmax_ = 1;
  npat_ = p_.isemicoarsen ? 3 : 1;

  // ---- level shapes of all patterns (the same for every grid of this size) ----
  dims_.assign(npat_, {});
  scs_.assign(npat_, {});
  for (int c = 0; c < npat_; c++) {
    dims_[c].push_back({n1, n2, n3});
    scs_[c].push_back({1, 1, 1});
    for (int lev = 1; lev < lmax_; lev++) {
      const auto& nf = dims_[c][lev - 1];
      std::array<int, 3> sc, n;
      for (int a = 0; a < 3; a++) sc[a] = (nf[a] % 2 == 0 && nf[a] > 2) ? 2 : 1;
      if (p_.isemicoarsen) sc[c] = 1;
      for (int a = 0; a < 3; a++) n[a] = nf[a] / sc[a];
      dims_[c].push_back(n);
      scs_[c].push_back(sc);
    }
  }

  meds_.resize(nmed);
  for (int m = 0; m < nmed; m++) {
    Medium& M = meds_[m];
    // ---- host geometry of all levels and patterns ----
    const float* xf[3] = {x1[m], x2[m], x3[m]};
    std::array<HostAxis, 3> a0;
    for (int a = 0; a < 3; a++) {
      std::vector<double> nodes(n_[a] + 1);
      for (int i = 0; i <= n_[a]; i++) nodes[i] = xf[a][i];
      a0[a].from_nodes(nodes);
    }
    M.axes.assign(npat_, {});
    for (int c = 0; c < npat_; c++) {
      M.axes[c].push_back(a0);
      for (int lev = 1; lev < lmax_; lev++) {
        const auto& n = dims_[c][lev];
        const auto& sc = scs_[c][lev];
        std::array<HostAxis, 3> ax;
        for (int a = 0; a < 3; a++) {
          std::vector<double> nodes(n[a] + 1);
          for (int i = 0; i <= n[a]; i++) nodes[i] = M.axes[c][lev - 1][a].x[sc[a] * i];
          ax[a].from_nodes(nodes);
        }
        M.axes[c].push_back(ax);
      }
    }

    // ---- medium: sigma*vol cells on the device, all levels and patterns ----
    const long ncell0 = (long)n1 * n2 * n3;
    std::vector<void*> tmp;
    std::vector<std::vector<double*>> cells(npat_ * lmax_);  // [pat*lmax+lev][3]
    double* W[3];
    for (int a = 0; a < 3; a++) {
      W[a] = (double*)dalloc(sizeof(double) * (n_[a] + 1), tmp);
      HIPCHECK(hipMemcpyAsync(W[a], a0[a].ds.data(), sizeof(double) * (n_[a] + 1), hipMemcpyHostToDevice, s_));
    }
    const float* sg[3] = {sig11[m], sig22[m], sig33[m]};
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

    if (p_.precision == 0) {
      build_hier<double>(M.h64, cells, lmax_, m);
      M.l0_64 = M.h64.pat[0][0];
    } else {
      build_hier<float>(M.h32, cells, lmax_, m);
      // FP64 level-0 operator of the refinement
      Hier<double> one;
      build_hier<double>(one, cells, 1, m);
      M.l0_64 = one.pat[0][0];
    }
    HIPCHECK(hipStreamSynchronize(s_));
    for (void* t : tmp) HIPCHECK(hipFree(t));
  }

  // records of the line solves: max over levels, patterns, axes and colours
  const size_t esize = p_.precision == 0 ? sizeof(LineScr<double>::type) : sizeof(LineScr<float>::type);
  for (int c = 0; c < npat_; c++)
    for (int lev = 0; lev < lmax_; lev++)
      for (int a = 0; a < 3; a++)
        for (int col = 0; col < 4; col++)
          lscr_bytes_ = std::max(lscr_bytes_, line_scratch_bytes(dims_[c][lev].data(), a, col, esize));

  slots_.resize(nslots);
  for (int k = 0; k < nslots; k++) {
    Slot& S = slots_[k];
    S.med = slot_med[k];
    S.omu = (double)omegas[k] * MU0;   // omega stored as float (src/gmg.c:1709)
    if (p_.precision == 0) {
      alloc_vectors<double>(S, meds_[S.med].h64);
      S.u64 = S.u[0];
      S.f64 = S.f[0];
      S.r64 = S.r[0];
    } else {
      alloc_vectors<float>(S, meds_[S.med].h32);
      const size_t b = sizeof(double2) * 3 * (size_t)nc0_;
      S.u64 = dalloc(b, owned_);
      S.f64 = dalloc(b, owned_);
      S.r64 = nullptr;   // scratch of blend_guess only, allocated there on first use
    }
    S.facs.assign((size_t)npat_ * lmax_ * 3, LineFactors());
    S.lscr = dalloc(std::max<size_t>(16, lscr_bytes_), owned_);
    S.partial = (double*)dalloc(sizeof(double) * residual_blocks_max(n_), owned_);
  }
  dnorm_ = (double*)dalloc(sizeof(double) * nslots, owned_);
  part_all_ = (double*)dalloc(sizeof(double) * MG_BATCH_MAX * residual_blocks_max(n_), owned_);
  HIPCHECK(hipHostMalloc((void**)&hnorm_, sizeof(double) * nslots, hipHostMallocDefault));
  HIPCHECK(hipStreamSynchronize(s_));
}

Solver::~Solver() {
  hipStreamSynchronize(s_);
  drop_graphs();
  for (void* p : owned_) hipFree(p);
  if (hnorm_) hipHostFree(hnorm_);
}

void Solver::set_omega(int slot, float omega) {
  if (slot < 0 || slot >= nslots()) throw std::runtime_error("slot out of range");
  slots_[slot].omu = (double)omega * MU0;
}

template <typename T>
void Solver::build_hier(Hier<T>& H, const std::vector<std::vector<double*>>& cells, int nlev, int med) {
  const auto& axes = meds_[med].axes;
  H.pat.assign(npat_, std::vector<DevLevel<T>>(nlev));
  H.maxnc.assign(nlev, 0);
  for (int c = 0; c < npat_; c++)
    for (int lev = 0; lev < nlev; lev++) {
      DevLevel<T>& L = H.pat[c][lev];
      const auto& n = dims_[c][lev];
      for (int a = 0; a < 3; a++) { L.n[a] = n[a]; L.sc[a] = scs_[c][lev][a]; }
      L.nc = (long)(n[0] + 1) * (n[1] + 1) * (n[2] + 1);
      if (L.nc > H.maxnc[lev]) H.maxnc[lev] = L.nc;
      // level 0 is shared by the patterns
      if (c > 0 && lev == 0) { L = H.pat[0][0]; continue; }
      const int nmax = std::max(n[0], std::max(n[1], n[2]));
      L.gstride = nmax + 1;
      // 1D geometry in FP64 for every storage precision (see Level)
      std::vector<double> g(15 * (size_t)L.gstride, 0.0);
      for (int a = 0; a < 3; a++) {
        const HostAxis& ax = axes[c][lev][a];
        for (int i = 0; i <= n[a]; i++) {
          g[(5 * a + 0) * L.gstride + i] = ax.ds[i];
          g[(5 * a + 1) * L.gstride + i] = 1.0 / ax.ds[i];
          g[(5 * a + 2) * L.gstride + i] = i > 0 ? 0.5 * (ax.ds[i] + ax.ds[i - 1]) : 0.0;
          g[(5 * a + 3) * L.gstride + i] = ax.x[i];
          g[(5 * a + 4) * L.gstride + i] = ax.d[i];
        }
      }
      L.geo = (double*)dalloc(sizeof(double) * g.size(), owned_);
      HIPCHECK(hipMemcpyAsync(L.geo, g.data(), sizeof(double) * g.size(), hipMemcpyHostToDevice, s_));
      if (lev > 0) {
        // weights of the finer level's nodes towards this level (src/gmg.c:1478-1488)
        const auto& nf = dims_[c][lev - 1];
        const int nfmax = std::max(nf[0], std::max(nf[1], nf[2]));
        L.wstride = nfmax + 1;
        std::vector<double> w(3 * (size_t)L.wstride, 0.0);
        for (int a = 0; a < 3; a++) {
          const HostAxis& fa = axes[c][lev - 1][a];
          const HostAxis& ca = axes[c][lev][a];
          for (int q = 0; q <= nf[a]; q++) {
            const int i = L.sc[a] == 2 ? q / 2 : q;
            if (i + 1 > n[a]) continue;
            w[a * L.wstride + q] = (fa.x[q] - ca.x[i]) / (ca.x[i + 1] - ca.x[i]);
          }
        }
        L.wr = (double*)dalloc(sizeof(double) * w.size(), owned_);
        HIPCHECK(hipMemcpyAsync(L.wr, w.data(), sizeof(double) * w.size(), hipMemcpyHostToDevice, s_));
      }
      L.sbar = (T*)dalloc(sizeof(T) * 3 * L.nc, owned_);
      const auto& cl = cells[c * lmax_ + lev];
      launch_edge_sigma<T>(n[0], n[1], n[2], cl[0], cl[1], cl[2], L.sbar, s_);
    }
}

template <typename T>
void Solver::alloc_vectors(Slot& S, const Hier<T>& H) {
  const int nlev = (int)H.maxnc.size();
  S.u.assign(nlev, nullptr);
  S.f.assign(nlev, nullptr);
  S.r.assign(nlev, nullptr);
  for (int lev = 0; lev < nlev; lev++) {
    const size_t bytes = sizeof(typename cvec<T>::type) * 3 * (size_t)H.maxnc[lev];
    S.u[lev] = dalloc(bytes, owned_);
    S.f[lev] = dalloc(bytes, owned_);
    S.r[lev] = dalloc(bytes, owned_);
    launch_zero<T>(S.u[lev], 3 * H.maxnc[lev], s_);
    launch_zero<T>(S.f[lev], 3 * H.maxnc[lev], s_);
    launch_zero<T>(S.r[lev], 3 * H.maxnc[lev], s_);
  }
}

// ---------------------------------------------------------------------------
// smoothing
// ---------------------------------------------------------------------------
template <typename T>
bool Solver::coarsens(int pat, int lev) {
  if (lev >= lmax_ - 1) return false;
  const auto& sc = scs_[pat][lev + 1];
  return sc[0] * sc[1] * sc[2] > 1;
}

// the coarsest level of a semi-coarsening pattern is two cells wide across
// the kept axis: its interior unknowns form a single line along that axis,
// so one line sweep is an exact solve
bool Solver::exact_coarsest(int pat) const {
  if (!p_.isemicoarsen) return false;
  const auto& n = dims_[pat][lmax_ - 1];
  return n[(pat + 1) % 3] == 2 && n[(pat + 2) % 3] == 2;
}

// factor records of (pattern, level, axis) for slot S, allocated on first use
// and rebuilt when the slot's frequency changed since they were made
template <typename T>
LineFactors& Solver::factors(Slot& S, int pat, int lev, int axis) {
  // records are sized for the cycle precision of this solver
  if ((sizeof(T) == sizeof(double)) != (p_.precision == 0))
    throw std::runtime_error("line solve precision differs from the solver's cycle precision");
  const int pe = lev == 0 ? 0 : pat;
  LineFactors& F = S.facs[((size_t)pe * lmax_ + lev) * 3 + axis];
  if (F.omu == S.omu) return F;
  const DevLevel<T>& L = hier<T>(S.med).pat[pe][lev];
  LinePack<T> P;
  P.njobs = 4;
  for (int col = 0; col < 4; col++) {
    if (!F.fac[col]) F.fac[col] = dalloc(std::max<size_t>(16, line_factor_bytes(L.n, axis, col, sizeof(typename LineFac<T>::type))), owned_);
    LineJob<T>& J = P.job[col];
    J.L = L.view(S.omu);
    J.u = nullptr;
    J.f = nullptr;
    J.fac = (typename LineFac<T>::type*)F.fac[col];
    J.scr = nullptr;
    line_job_colour(J, axis, col);
  }
  launch_line_factor<T>(P, axis, s_);
  F.omu = S.omu;
  return F;
}

template <typename T>
void Solver::line_sweep_on(int pat, int lev, int axis, int iter, void* u, const void* f) {
  Slot& S = slots_[0];
  const LineFactors& F = factors<T>(S, pat, lev, axis);
  const DevLevel<T>& L = hier<T>(S.med).pat[pat][lev];
  LinePack<T> P;
  P.njobs = 1;
  LineJob<T>& J = P.job[0];
  J.L = L.view(S.omu);
  J.u = (typename cvec<T>::type*)u;
  J.f = (const typename cvec<T>::type*)f;
  J.scr = (typename LineScr<T>::type*)S.lscr;
  for (int k = 0; k < 4; k++) {
    const int colour = (iter % 2 == 1) ? 3 - k : k;
    J.fac = (typename LineFac<T>::type*)F.fac[colour];
    line_job_colour(J, axis, colour);
    launch_line_solve<T>(P, axis, s_);
  }
}

template void Solver::line_sweep_on<double>(int, int, int, int, void*, const void*);
template void Solver::line_sweep_on<float>(int, int, int, int, void*, const void*);

// one sweep along `axis` for every active slot: per colour, the slots share
// launches of up to LINE_PACK_MAX jobs
template <typename T>
void Solver::line_sweep(int pat, int lev, int axis, int iter) {
  std::vector<const LineFactors*> F(act_.size());
  for (size_t i = 0; i < act_.size(); i++) F[i] = &factors<T>(slots_[act_[i]], pat, lev, axis);
  for (int k = 0; k < 4; k++) {
    const int colour = (iter % 2 == 1) ? 3 - k : k;
    for (size_t i0 = 0; i0 < act_.size(); i0 += LINE_PACK_MAX) {
      LinePack<T> P;
      P.njobs = (int)std::min<size_t>(LINE_PACK_MAX, act_.size() - i0);
      for (int j = 0; j < P.njobs; j++) {
        Slot& S = slots_[act_[i0 + j]];
        LineJob<T>& J = P.job[j];
        J.L = hier<T>(S.med).pat[pat][lev].view(S.omu);
        J.u = (typename cvec<T>::type*)S.u[lev];
        J.f = (const typename cvec<T>::type*)S.f[lev];
        J.fac = (typename LineFac<T>::type*)F[i0 + j]->fac[colour];
        J.scr = (typename LineScr<T>::type*)S.lscr;
        line_job_colour(J, axis, colour);
      }
      launch_line_solve<T>(P, axis, s_);
    }
  }
}

// one vertex-patch Gauss-Seidel sweep for every active slot
template <typename T>
void Solver::point_sweep(int pat, int lev, int iter) {
  for (size_t i0 = 0; i0 < act_.size(); i0 += MG_BATCH_MAX) {
    const int n = (int)std::min<size_t>(MG_BATCH_MAX, act_.size() - i0);
    void* u[MG_BATCH_MAX];
    const void* f[MG_BATCH_MAX];
    double omu[MG_BATCH_MAX];
    Level<T> Ls[MG_BATCH_MAX];
    for (int j = 0; j < n; j++) {
      Slot& S = slots_[act_[i0 + j]];
      u[j] = S.u[lev];
      f[j] = S.f[lev];
      omu[j] = S.omu;
      Ls[j] = hier<T>(S.med).pat[pat][lev].view(S.omu);
    }
    launch_point_gs_batch<T>(Ls, n, u, f, omu, iter, s_);
  }
}

// src/gmg.c:1301-1313: line sweeps along the axes coarsened on the next
// level with semi-coarsening; the vertex-patch point smoother everywhere
// without it.  On the coarsest level of a semi-coarsening pattern the lines
// along the kept axis replace the reference's point smoother (one line
// sweep there is an exact solve, see exact_coarsest).
template <typename T>
void Solver::smooth(int pat, int lev, int iter) {
  if (p_.isemicoarsen && lev < lmax_ - 1) {
    const auto& sc = scs_[pat][lev + 1];
    for (int a = 0; a < 3; a++)
      if (sc[a] == 2) line_sweep<T>(pat, lev, a, iter);
  } else if (p_.isemicoarsen) {
    for (int s = 0; s < p_.coarse_sweeps; s++) line_sweep<T>(pat, lev, pat, iter + s);
  } else {
    point_sweep<T>(pat, lev, iter);
  }
}

// ---------------------------------------------------------------------------
// autotuning of the line-sweep launches
// ---------------------------------------------------------------------------
template <typename T>
int Solver::autotune_t(int reps) {
  if (reps < 1) reps = 1;
  std::vector<int> all(slots_.size());
  for (size_t k = 0; k < slots_.size(); k++) all[k] = (int)k;
  hipEvent_t e0, e1;
  HIPCHECK(hipEventCreate(&e0));
  HIPCHECK(hipEventCreate(&e1));
  int tuned = 0;
  // microseconds per launch: the best of three rounds of reps launches
  // (launches of a few microseconds vary by more than the choices differ)
  auto time_of = [&](const LinePack<T>& P, int axis, int choice) {
    launch_line_solve_as<T>(P, axis, choice, s_);   // warm (instruction cache, TLB)
    double best = 0;
    for (int round = 0; round < 3; round++) {
      HIPCHECK(hipEventRecord(e0, s_));
      for (int r = 0; r < reps; r++) launch_line_solve_as<T>(P, axis, choice, s_);
      HIPCHECK(hipEventRecord(e1, s_));
      HIPCHECK(hipEventSynchronize(e1));
      float ms = 0;
      HIPCHECK(hipEventElapsedTime(&ms, e0, e1));
      const double us = 1e3 * (double)ms / reps;
      if (round == 0 || us < best) best = us;
    }
    return best;
  };
  for (int pat = 0; pat < npat_; pat++) {
    for (int lev = 0; lev + 1 < lmax_ && p_.isemicoarsen; lev++) {
      for (int axis = 0; axis < 3; axis++) {
        if (scs_[pat][lev + 1][axis] != 2) continue;
        std::vector<const LineFactors*> F(all.size());
        for (size_t i = 0; i < all.size(); i++) F[i] = &factors<T>(slots_[all[i]], pat, lev, axis);
        for (int colour = 0; colour < 4; colour++) {
          for (size_t i0 = 0; i0 < all.size(); i0 += LINE_PACK_MAX) {
            LinePack<T> P;
            P.njobs = (int)std::min<size_t>(LINE_PACK_MAX, all.size() - i0);
            for (int j = 0; j < P.njobs; j++) {
              Slot& S = slots_[all[i0 + j]];
              LineJob<T>& J = P.job[j];
              J.L = hier<T>(S.med).pat[pat][lev].view(S.omu);
              J.u = (typename cvec<T>::type*)S.u[lev];
              J.f = (const typename cvec<T>::type*)S.f[lev];
              J.fac = (typename LineFac<T>::type*)F[i0 + j]->fac[colour];
              J.scr = (typename LineScr<T>::type*)S.lscr;
              line_job_colour(J, axis, colour);
            }
            if (line_pcr_shape<T>(P, axis)) continue;
            const LineShape key = line_shape<T>(P, axis);
            if (key.nl < 1 || has_line_choice(key)) continue;
            double best = 0;
            int bc = LINE_RULE;
            for (int c = 0; c < LINE_NCHOICE; c++) {
              const double us = time_of(P, axis, c);
              if (bc == LINE_RULE || us < best) { best = us; bc = c; }
            }
            set_line_choice(key, bc, best, time_of(P, axis, LINE_RULE_ONLY));
            tuned++;
          }
        }
      }
    }
  }
  HIPCHECK(hipEventDestroy(e0));
  HIPCHECK(hipEventDestroy(e1));
  return tuned;
}

void Solver::blend_guess(int n, const void* const* b, const void* const* x1, void* const* xg, double* beta) {
  if (n < 0 || n > (int)slots_.size()) throw std::runtime_error("blend_guess: slot count out of range");
  // b, x1 and xg may come from kernels on other streams
  HIPCHECK(hipDeviceSynchronize());
  (void)hipGetLastError();   // as in solve_batch: the final check reports this call's launches
  const long nv = 3 * nc0_;
  if (!blend_part_) {
    blend_part_ = (double*)dalloc(sizeof(double) * 3 * (size_t)blend_blocks(nv), owned_);
    blend_dots_ = (double*)dalloc(sizeof(double) * 3, owned_);
  }
  double *part = blend_part_, *dots = blend_dots_;
  for (int k = 0; k < n; k++) {
    Slot& S = slots_[k];
    if (!S.r64) S.r64 = dalloc(sizeof(double2) * nv, owned_);
    const DevLevel<double>& L0 = meds_[S.med].l0_64;
    launch_scale_rhs<double>(L0.view(S.omu), L0.D(0), L0.D(1), L0.D(2), b[k], S.f64, s_);
    // residuals of the iterate (r64) and of the guess (u64, free before a solve)
    launch_residual<double>(L0.view(S.omu), x1[k], S.f64, S.r64, nullptr, s_);
    launch_residual<double>(L0.view(S.omu), xg[k], S.f64, S.u64, nullptr, s_);
    launch_blend_dots(S.r64, S.u64, nv, part, dots, s_);
    double h[3];
    HIPCHECK(hipMemcpyAsync(h, dots, sizeof(h), hipMemcpyDeviceToHost, s_));
    HIPCHECK(hipStreamSynchronize(s_));
    double br = 0, bi = 0;
    if (std::isfinite(h[0]) && std::isfinite(h[1]) && std::isfinite(h[2]) && h[2] > 0) {
      br = h[0] / h[2];
      bi = h[1] / h[2];
    }
    launch_blend(xg[k], x1[k], br, bi, nv, s_);
    if (beta) {
      beta[2 * k] = br;
      beta[2 * k + 1] = bi;
    }
  }
  HIPCHECK(hipStreamSynchronize(s_));
  HIPCHECK(hipGetLastError());
}

int Solver::autotune(int reps) {
  HIPCHECK(hipDeviceSynchronize());
  drop_graphs();
  const int n = p_.precision == 0 ? autotune_t<double>(reps) : autotune_t<float>(reps);
  HIPCHECK(hipStreamSynchronize(s_));
  return n;
}

// ---------------------------------------------------------------------------
// kernel rates: the level-0 kernels of slot 0 timed one by one
// ---------------------------------------------------------------------------
template <typename T>
void Solver::kernel_rates_t(int reps, double* ms, double* bytes) {
  Slot& S = slots_[0];
  const DevLevel<T>& L = hier<T>(S.med).pat[0][0];
  const DevLevel<T>& C = hier<T>(S.med).pat[0][1];
  const Level<T> V = L.view(S.omu), VC = C.view(S.omu);
  const Transfer<T> X = C.transfer();
  const double es = sizeof(typename cvec<T>::type), cs = sizeof(T);
  const double ne = 3.0 * (double)L.nc, nc = (double)L.n[0] * L.n[1] * L.n[2], nec = 3.0 * (double)C.nc;
  const double sweep = 3.0 * ne * es + 3.0 * nc * cs;
  hipEvent_t e0, e1;
  HIPCHECK(hipEventCreate(&e0));
  HIPCHECK(hipEventCreate(&e1));
  for (int a = 0; a < 3; a++) factors<T>(S, 0, 0, a);
  auto timed = [&](int k, double by, auto&& launch) {
    launch();   // warm
    HIPCHECK(hipEventRecord(e0, s_));
    for (int r = 0; r < reps; r++) launch();
    HIPCHECK(hipEventRecord(e1, s_));
    HIPCHECK(hipEventSynchronize(e1));
    float t = 0;
    HIPCHECK(hipEventElapsedTime(&t, e0, e1));
    ms[k] = t / reps;
    bytes[k] = by;
  };
  timed(0, ne * (3.0 * es) + 3.0 * nc * cs, [&] { launch_residual<T>(V, S.u[0], S.f[0], S.r[0], S.partial, s_); });
  // restriction: fine residual read, coarse right-hand side and the zeroed
  // coarse correction written
  timed(1, (ne + 2.0 * nec) * es, [&] { launch_restrict<T>(V, VC, X, S.r[0], S.f[1], S.u[1], s_); });
  timed(2, (2.0 * ne + nec) * es, [&] { launch_prolong<T>(V, VC, X, S.u[1], S.u[0], s_); });
  for (int a = 0; a < 3; a++)
    timed(3 + a, sweep, [&] { line_sweep_on<T>(0, 0, a, 0, S.u[0], S.f[0]); });
  HIPCHECK(hipEventDestroy(e0));
  HIPCHECK(hipEventDestroy(e1));
}

int Solver::kernel_rates(int reps, double* ms, double* bytes) {
  if (reps < 1) throw std::runtime_error("kernel_rates needs reps >= 1");
  if (lmax_ < 2) throw std::runtime_error("kernel_rates needs a coarse level");
  HIPCHECK(hipDeviceSynchronize());
  if (p_.precision == 0) kernel_rates_t<double>(reps, ms, bytes);
  else kernel_rates_t<float>(reps, ms, bytes);
  HIPCHECK(hipStreamSynchronize(s_));
  return 6;
}

// ---------------------------------------------------------------------------
// convergence tests
// ---------------------------------------------------------------------------
void Solver::norms(const std::vector<int>& which, bool add, std::vector<double>& out) {
  const int nb = residual_blocks<double>(n_);
  for (size_t i0 = 0; i0 < which.size(); i0 += MG_BATCH_MAX) {
    const int n = (int)std::min<size_t>(MG_BATCH_MAX, which.size() - i0);
    const void* u[MG_BATCH_MAX];
    const void* f[MG_BATCH_MAX];
    void* r[MG_BATCH_MAX];
    void* r32[MG_BATCH_MAX];
    double omu[MG_BATCH_MAX];
    Level<double> Ls[MG_BATCH_MAX];
    for (int j = 0; j < n; j++) {
      Slot& S = slots_[which[i0 + j]];
      Ls[j] = meds_[S.med].l0_64.view(S.omu);
      u[j] = S.u64;
      f[j] = S.f64;
      r[j] = S.r64;
      r32[j] = S.f[0];
      omu[j] = S.omu;
      if (add) launch_promote_add(S.u64, S.u[0], 3 * nc0_, s_);
    }
    if (p_.precision == 0) launch_residual_batch<double>(Ls, n, u, f, r, omu, s_, part_all_);
    else launch_refine_test(Ls, n, u, f, r32, part_all_, s_);
    launch_sum_partials_batch(part_all_, nb, n, dnorm_ + i0, s_);
  }
  HIPCHECK(hipMemcpyAsync(hnorm_, dnorm_, sizeof(double) * which.size(), hipMemcpyDeviceToHost, s_));
  HIPCHECK(hipStreamSynchronize(s_));
  out.resize(which.size());
  for (size_t i = 0; i < which.size(); i++) out[i] = std::sqrt(hnorm_[i]);
}

bool Solver::check_active(int pat) {
  (void)pat;
  std::vector<double> rn;
  norms(act_, pending_, rn);
  pending_ = false;
  std::vector<int> keep;
  for (size_t i = 0; i < act_.size(); i++) {
    Slot& S = slots_[act_[i]];
    SolveStats* st = S.st;
    st->history.push_back(rn[i]);
    if (p_.verb) printf("slot=%d icycle=%d rnorm=%e\n", act_[i], icycle_, rn[i]);
    if (!std::isfinite(rn[i])) {
      st->status = 2;
      st->cycles = icycle_;
      continue;
    }
    st->rnorm = rn[i];
    if (icycle_ == 0 && rnorm0_in_cycle_) {
      // FP64 V- and W-cycles: the reference takes rnorm0 from the first
      // in-cycle norm, after the pre-smoothing of cycle 0, and tests from
      // cycle 1 on (src/gmg.c:1594-1599, 1651-1656)
      S.rnorm0 = rn[i];
      st->rnorm0 = rn[i];
      keep.push_back(act_[i]);
      continue;
    }
    if (rn[i] < S.rnorm0 * p_.tol) {
      st->cycles = icycle_;
      continue;
    }
    keep.push_back(act_[i]);
  }
  act_.swap(keep);
  return act_.empty();
}

// ---------------------------------------------------------------------------
// cycles
// ---------------------------------------------------------------------------
// restriction of the level's residual (or right-hand side when the guess is
// zero) followed by the recursive coarse solve and the prolongation; returns
// true when the level-0 convergence test retired every slot
template <typename T>
bool Solver::coarse_correction(int pat, int lev, bool norm_here, bool zero_guess, int kind) {
  // level shapes (the same for every slot's grid)
  const DevLevel<T>& L = hier<T>(0).pat[pat][lev];
  const DevLevel<T>& C = hier<T>(0).pat[pat][lev + 1];
  bool from_f = false, r0_was_ready = false;
  if (lev == 0 && norm_here) {
    // level-0 residual with its norm (FP64 direct form: r64 is the level's r)
    if (check_active(pat)) return true;
  } else if (lev == 0 && r0_ready_) {
    // the cycle's convergence test has just computed r[0] = f - A u
    r0_ready_ = false;
    r0_was_ready = true;
  } else if (zero_guess) {
    from_f = true;
  }
  // the active slots share the level: one launch per step for up to
  // MG_BATCH_MAX of them (grid y = slot)
  const bool resid = !(lev == 0 && (norm_here || r0_was_ready)) && !from_f;
  for (size_t i0 = 0; i0 < act_.size(); i0 += MG_BATCH_MAX) {
    const int n = (int)std::min<size_t>(MG_BATCH_MAX, act_.size() - i0);
    const void* u[MG_BATCH_MAX];
    const void* f[MG_BATCH_MAX];
    const void* rf[MG_BATCH_MAX];
    void* r[MG_BATCH_MAX];
    void* fc[MG_BATCH_MAX];
    void* uc[MG_BATCH_MAX];
    double omu[MG_BATCH_MAX];
    Level<T> Ls[MG_BATCH_MAX];
    Transfer<T> Xs[MG_BATCH_MAX];
    for (int j = 0; j < n; j++) {
      Slot& S = slots_[act_[i0 + j]];
      u[j] = S.u[lev];
      f[j] = S.f[lev];
      r[j] = S.r[lev];
      rf[j] = from_f ? S.f[lev] : S.r[lev];
      fc[j] = S.f[lev + 1];
      uc[j] = S.u[lev + 1];
      omu[j] = S.omu;
      Ls[j] = hier<T>(S.med).pat[pat][lev].view(S.omu);
      Xs[j] = hier<T>(S.med).pat[pat][lev + 1].transfer();
    }
    if (resid) launch_residual_batch<T>(Ls, n, u, f, r, omu, s_);
    launch_restrict_batch<T>(L.view(0.0), C.view(0.0), Xs, n, rf, fc, uc, s_);
  }
  if (kind == 2) f_cycle<T>(pat, lev + 1, false, true);
  else if (kind == 3) { w_cycle<T>(pat, lev + 1, false, true); w_cycle<T>(pat, lev + 1, false, false); }
  else v_cycle<T>(pat, lev + 1, false, true);
  for (size_t i0 = 0; i0 < act_.size(); i0 += MG_BATCH_MAX) {
    const int n = (int)std::min<size_t>(MG_BATCH_MAX, act_.size() - i0);
    const void* uc[MG_BATCH_MAX];
    void* uf[MG_BATCH_MAX];
    Transfer<T> Xs[MG_BATCH_MAX];
    for (int j = 0; j < n; j++) {
      Slot& S = slots_[act_[i0 + j]];
      uc[j] = S.u[lev + 1];
      uf[j] = S.u[lev];
      Xs[j] = hier<T>(S.med).pat[pat][lev + 1].transfer();
    }
    launch_prolong_batch<T>(L.view(0.0), C.view(0.0), Xs, n, uc, uf, s_);
  }
  return false;
}

// src/gmg.c:1586-1612
template <typename T>
bool Solver::v_cycle(int pat, int lev, bool norm_here, bool zero_guess) {
  if (lev == lmax_ - 1 && lev > 0 && exact_coarsest(pat)) {
    line_sweep<T>(pat, lev, pat, 0);
    return false;
  }
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
  if (lev == lmax_ - 1) {
    for (int k : act_) launch_zero<T>(slots_[k].u[lev], 3 * hier<T>(slots_[k].med).pat[pat][lev].nc, s_);
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
  if (lev == lmax_ - 1 && lev > 0 && exact_coarsest(pat)) {
    line_sweep<T>(pat, lev, pat, 0);
    return false;
  }
  for (int i = 0; i < p_.v1; i++) { smooth<T>(pat, lev, i); zero_guess = false; }
  if (coarsens<T>(pat, lev)) {
    if (coarse_correction<T>(pat, lev, norm_here && p_.cycleopt == 3, zero_guess, 3)) return true;
  }
  for (int i = 0; i < p_.v2; i++) smooth<T>(pat, lev, i);
  return false;
}

// ---------------------------------------------------------------------------
// cycle bodies and their graphs
// ---------------------------------------------------------------------------
void Solver::cycle_body_eager(int pat) {
  const int kind = p_.cycleopt;
  const long nv = 3 * nc0_;
  if (p_.precision == 0) {
    // F-cycle after its opening test, which left r[0] = f - A u
    r0_ready_ = true;
    f_cycle<double>(pat, 0, false, false);
    r0_ready_ = false;
    return;
  }
  // refinement: the test left f[0] = f - A u64 in FP32; FP32-storage cycle
  // on it from zero.  Its correction u[0] is added to u64 by the next test
  // (or after the last cycle).
  for (int k : act_) launch_zero<float>(slots_[k].u[0], nv, s_);
  if (kind == 1) v_cycle<float>(pat, 0, false, true);
  else if (kind == 3) w_cycle<float>(pat, 0, false, true);
  else f_cycle<float>(pat, 0, false, true);
}

void Solver::drop_graphs() {
  for (auto& g : graphs_) hipGraphExecDestroy(g.second);
  graphs_.clear();
}

// The sequence of launches of a cycle body depends only on the pattern, the
// active slots and their frequencies (kernel arguments carry omega mu0), so
// the body is captured once per (pattern, active set) and replayed.  The
// first cycle of a pattern at a new frequency runs eagerly: it builds the
// line factors, which must not be captured.
void Solver::cycle_body(int pat) {
  if (!p_.graphs || timing::enabled() || slots_.size() > (size_t)MAX_SLOTS) {
    cycle_body_eager(pat);
    return;
  }
  bool fresh = false;
  for (int k : act_)
    if (slots_[k].seen_omu[pat] != slots_[k].omu) fresh = true;
  if (fresh) {
    cycle_body_eager(pat);
    for (int k : act_) slots_[k].seen_omu[pat] = slots_[k].omu;
    return;
  }
  std::vector<double> om(slots_.size());
  for (size_t k = 0; k < slots_.size(); k++) om[k] = slots_[k].omu;
  if (om != graph_omu_) {
    drop_graphs();
    graph_omu_ = om;
  }
  unsigned long long mask = 0;
  for (int k : act_) mask |= 1ull << k;
  const auto key = std::make_pair(pat, mask);
  auto it = graphs_.find(key);
  if (it == graphs_.end()) {
    hipGraph_t g = nullptr;
    hipGraphExec_t ge = nullptr;
    HIPCHECK(hipStreamBeginCapture(s_, hipStreamCaptureModeThreadLocal));
    cycle_body_eager(pat);
    HIPCHECK(hipStreamEndCapture(s_, &g));
    // kernel launches report their errors here, not by a return code
    check_hip(hipGetLastError(), "a kernel launch while capturing a cycle body");
    HIPCHECK(hipGraphInstantiate(&ge, g, nullptr, nullptr, 0));
    HIPCHECK(hipGraphDestroy(g));
    it = graphs_.emplace(key, ge).first;
  }
  HIPCHECK(hipGraphLaunch(it->second, s_));
}

// ---------------------------------------------------------------------------
// solve (src/gmg.c:1894-1932)
// ---------------------------------------------------------------------------
int Solver::solve(const void* b, void* x, void* h, const void* x0, SolveStats* st) {
  SolveStats local;
  return solve_batch(1, &b, &x, &h, &x0, st ? st : &local);
}

int Solver::solve_batch(int n, const void* const* b, void* const* x, void* const* h, const void* const* x0,
                        SolveStats* st) {
  if (n < 1 || n > nslots()) throw std::runtime_error("batch size out of range of the solver's slots");
  const long nv = 3 * nc0_;
  // The launch-error check at the end of the solve reads the thread's last
  // HIP error; clear what calls outside the solver left there (unchecked
  // frees and queries of other objects) so that it reports this solve's.
  (void)hipGetLastError();
  // The solver's stream does not order itself after other streams: wait for
  // the work the caller queued elsewhere (b and x0 may still be in flight).
  HIPCHECK(hipDeviceSynchronize());
  act_.clear();
  for (int k = 0; k < n; k++) {
    Slot& S = slots_[k];
    S.st = &st[k];
    *S.st = SolveStats();
    S.st->cycles = -1;
    // f = b * dual volume (src/gmg.c:1899-1910)
    const DevLevel<double>& L0 = meds_[S.med].l0_64;
    launch_scale_rhs<double>(L0.view(S.omu), L0.D(0), L0.D(1), L0.D(2), b[k], S.f64, s_);
    launch_zero<double>(S.u64, nv, s_);
    act_.push_back(k);
  }
  pending_ = false;
  // ||f||: the reference's rnorm0 (its first residual, taken at u = 0)
  {
    std::vector<double> rn;
    norms(act_, false, rn);
    for (size_t i = 0; i < act_.size(); i++) {
      slots_[act_[i]].rnorm0 = rn[i];
      slots_[act_[i]].st->rnorm0 = rn[i];
    }
  }
  // Initial guesses are judged before any cycle: a guess whose residual
  // ||f - A x0|| is not below the zero guess's ||f|| (or is not finite) is
  // dropped and the slot starts from zero.  The stopping test stays relative
  // to ||f|| either way.
  {
    std::vector<int> g;
    for (int k = 0; k < n; k++)
      if (x0 && x0[k]) {
        HIPCHECK(hipMemcpyAsync(slots_[k].u64, x0[k], sizeof(double2) * nv, hipMemcpyDeviceToDevice, s_));
        g.push_back(k);
      }
    if (!g.empty()) {
      std::vector<double> rn;
      norms(g, false, rn);
      for (size_t i = 0; i < g.size(); i++) {
        Slot& S = slots_[g[i]];
        S.st->x0_rnorm = rn[i];
        if (!(rn[i] <= S.rnorm0)) {
          launch_zero<double>(S.u64, nv, s_);
          S.st->x0_rejected = 1;
        }
      }
    }
  }

  const int kind = p_.cycleopt;
  rnorm0_in_cycle_ = p_.precision == 0 && kind != 2;
  for (icycle_ = 0; icycle_ < p_.ncycle && !act_.empty(); icycle_++) {
    const int pat = p_.isemicoarsen ? (icycle_ + first_cycle_) % 3 : 0;
    const bool test_first = p_.precision != 0 || (kind == 2 && coarsens<double>(pat, 0));
    if (test_first) {
      // the convergence test opens the cycle (src/gmg.c:1629-1630 for the
      // F-cycle); the rest of the cycle runs as one body
      if (check_active(pat)) break;
      cycle_body(pat);
      pending_ = p_.precision != 0;
      continue;
    }
    // FP64 V/W cycles test after the pre-smoothing (src/gmg.c:1599,1656)
    if (kind == 1) v_cycle<double>(pat, 0, true, false);
    else if (kind == 3) w_cycle<double>(pat, 0, true, false);
    else f_cycle<double>(pat, 0, true, false);
    if (act_.empty()) break;
    if (lmax_ == 1 || !coarsens<double>(pat, 0)) {
      // no coarse level carries the norm test: test after the cycle
      icycle_++;
      const bool done = check_active(pat);
      icycle_--;
      if (done) break;
    }
  }
  // slots that ran out of cycles still hold their last cycle's correction
  if (pending_)
    for (int k : act_) launch_promote_add(slots_[k].u64, slots_[k].u[0], nv, s_);
  pending_ = false;
  for (int k = 0; k < n; k++) {
    Slot& S = slots_[k];
    SolveStats* s = S.st;
    if (s->cycles < 0) s->cycles = p_.ncycle;
    if (s->status == 0 && !(s->rnorm < S.rnorm0 * p_.tol)) s->status = 1;
    if (s->history.empty()) s->status = 1;
    HIPCHECK(hipMemcpyAsync(x[k], S.u64, sizeof(double2) * nv, hipMemcpyDeviceToDevice, s_));
    launch_curl_h<double>(meds_[S.med].l0_64.view(S.omu), S.u64, h[k], s_);
  }
  HIPCHECK(hipStreamSynchronize(s_));
  HIPCHECK(hipGetLastError());
  for (int k = 0; k < n; k++) slots_[k].st = nullptr;
  act_.clear();
  return 0;
}

}  // namespace emmi
