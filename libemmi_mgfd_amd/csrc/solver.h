// solver.h -- multigrid Maxwell solver on one GPU (host orchestration).
#pragma once

#include <hip/hip_runtime.h>
#include <array>
#include <vector>

#include "emmi_dev.h"
#include "mg_kernels.h"
#include "line_kernels.h"

namespace emmi {

struct SolveParams {
  int cycleopt = 2;      // 1 V, 2 F, 3 W (src/gmg.c:1712)
  int ncycle = 30;
  int v1 = 1, v2 = 1;
  int isemicoarsen = 1;
  double tol = 1e-6;
  int precision = 0;     // 0: complex FP64 throughout; 1: FP32 cycles inside FP64 refinement
  int coarse_sweeps = 2; // line sweeps per smoothing call on the coarsest level
  int verb = 0;
};

struct SolveStats {
  int cycles = 0;
  int status = 0;        // 0 converged, 1 not converged, 2 non-finite residual
  double rnorm0 = 0, rnorm = 0;
  std::vector<double> history;
};

void check_hip(hipError_t e, const char* what);

struct HostAxis {
  std::vector<double> x, xs, d, ds;
  void from_nodes(const std::vector<double>& nodes);  // src/gmg.c:1672-1685
};

// geometry and medium of one level in working precision T
template <typename T>
struct DevLevel {
  int n[3] = {0, 0, 0};
  int sc[3] = {1, 1, 1};
  long nc = 0;
  T* geo = nullptr;   // per axis a, at offset 5*a*(nmax+1): W, IW, H, X, D (n_a+1 each)
  int gstride = 0;
  T* wr = nullptr;    // weights of the finer level towards this one (3 x finer nmax+1)
  int wstride = 0;
  T* sbar = nullptr;  // 3*nc
  Level<T> view(double omu) const;
  Transfer<T> transfer() const;
  const T* D(int a) const { return geo + (5 * a + 4) * gstride; }
};

// factored line systems of one (level, axis), per colour, for one frequency
struct LineFactors {
  double* fac[4] = {nullptr, nullptr, nullptr, nullptr};
  double omu = -1.0;   // frequency the records were built for
};

template <typename T>
struct Hier {
  std::vector<std::vector<DevLevel<T>>> pat;  // [pattern][level]
  std::vector<void*> u, f, r;                 // per level, sized to the max over patterns
};

class Solver {
 public:
  // x1..x3: float node coordinates (n+1); sig: float [n3][n2][n1] conductivities
  // (host pointers when sig_on_device == 0, device pointers otherwise)
  Solver(int n1, int n2, int n3, const float* x1, const float* x2, const float* x3, const float* sig11,
         const float* sig22, const float* sig33, int sig_on_device, float omega, const SolveParams& p,
         hipStream_t stream);
  ~Solver();
  Solver(const Solver&) = delete;
  Solver& operator=(const Solver&) = delete;

  // b, x, h: device vectors of 3(n1+1)(n2+1)(n3+1) complex doubles.  b is the
  // source term (before the dual-volume scaling); x receives E, h receives H
  // (h may alias b).  x0 (device, may be null) is an initial guess.
  int solve(const void* b, void* x, void* h, const void* x0, SolveStats* st);

  long nvec() const { return 3 * nc0_; }
  int lmax() const { return lmax_; }
  int npatterns() const { return npat_; }
  double omu() const { return omu_; }
  // the hierarchy depends on the grid and medium only: switch frequency
  void set_omega(float omega) { omu_ = (double)omega * (4. * 3.14159265358979323846264 * 1e-7); }
  const int* dims(int pat, int lev) const { return dims_[pat][lev].data(); }

  // kernel-level access for the element-wise checks (FP64 hierarchy)
  const Hier<double>& h64() const { return h64_; }
  const Hier<float>& h32() const { return h32_; }
  hipStream_t stream() const { return s_; }
  // one line sweep (all four colours) along `axis` on vectors u, f of level
  // `lev` of pattern `pat`, in the working precision T of the solver
  template <typename T> void line_sweep_on(int pat, int lev, int axis, int iter, void* u, const void* f);

 private:
  template <typename T> Hier<T>& hier();
  template <typename T> void build_hier(Hier<T>& H, const std::vector<std::vector<double*>>& cells, int nlev);
  template <typename T> void smooth(int pat, int lev, int iter);
  template <typename T> void line_sweep(int pat, int lev, int axis, int iter);
  template <typename T> bool coarsens(int pat, int lev);
  template <typename T> LineFactors& factors(int pat, int lev, int axis);
  template <typename T> bool v_cycle(int pat, int lev, bool norm_here, bool zero_guess);
  template <typename T> bool f_cycle(int pat, int lev, bool norm_here, bool zero_guess);
  template <typename T> bool w_cycle(int pat, int lev, bool norm_here, bool zero_guess);
  template <typename T> bool coarse_correction(int pat, int lev, bool norm_here, bool zero_guess, int kind);
  bool check(double rn);
  double norm_of(const DevLevel<double>& L, const void* u, const void* f, void* r);
  double vec_norm(const void* v);

  int n_[3];
  long nc0_;
  int lmax_, npat_;
  double omu_;
  SolveParams p_;
  hipStream_t s_;
  std::vector<std::vector<std::array<int, 3>>> dims_, scs_;   // [pattern][level]
  std::vector<std::vector<std::array<HostAxis, 3>>> axes_;    // [pattern][level][axis]
  Hier<double> h64_;
  Hier<float> h32_;
  DevLevel<double> l0_64_;    // level-0 FP64 operator of the refinement
  std::vector<LineFactors> facs_;   // [(pattern*lmax + level)*3 + axis]; level 0 under pattern 0
  double* lscr_ = nullptr;          // substitution scratch of the line solves
  void* u64_ = nullptr;
  void* f64_ = nullptr;
  void* r64_ = nullptr;
  double* partial_ = nullptr;
  double* dnorm_ = nullptr;
  double* hnorm_ = nullptr;   // pinned host
  std::vector<void*> owned_;
  // convergence state of the running solve
  int icycle_ = 0;
  double rnorm0_ = 0;
  SolveStats* st_ = nullptr;
  bool stopped_ = false;
};

}  // namespace emmi
