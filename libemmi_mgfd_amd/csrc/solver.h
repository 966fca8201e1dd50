// solver.h -- multigrid Maxwell solver on one GPU (host orchestration).
//
// One Solver owns the level hierarchy of one grid and medium (shared by all
// frequencies of a shot) and a number of right-hand-side slots, each with its
// own frequency, level vectors, factored line systems and convergence state.
// solve_batch runs the slots in lockstep -- every kernel launch of a cycle
// serves all active slots -- and retires each slot when its residual test
// passes, so every slot follows exactly the iteration it would follow alone.
#pragma once

#include <hip/hip_runtime.h>
#include <array>
#include <map>
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
  int graphs = 1;        // replay cycle bodies as captured HIP graphs
};

struct SolveStats {
  int cycles = 0;
  int status = 0;        // 0 converged, 1 not converged, 2 non-finite residual
  double rnorm0 = 0, rnorm = 0;
  int x0_rejected = 0;   // the initial guess lost to the zero guess and was dropped
  double x0_rnorm = 0;   // ||f - A x0|| of the initial guess (0 without one)
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
  double* geo = nullptr;   // per axis a, at offset 5*a*(nmax+1): W, IW, H, X, D (n_a+1 each)
  int gstride = 0;
  double* wr = nullptr;    // weights of the finer level towards this one (3 x finer nmax+1)
  int wstride = 0;
  T* sbar = nullptr;  // 3*nc
  Level<T> view(double omu) const;
  Transfer<T> transfer() const;
  const double* D(int a) const { return geo + (5 * a + 4) * gstride; }
};

template <typename T>
struct Hier {
  std::vector<std::vector<DevLevel<T>>> pat;  // [pattern][level]
  std::vector<long> maxnc;                    // per level, max over patterns
};

// factored line systems of one (level, axis), per colour, for one frequency
struct LineFactors {
  void* fac[4] = {nullptr, nullptr, nullptr, nullptr};  // values of the cycle precision
  double omu = -1.0;   // frequency the records were built for
};

// grid and medium of one shot: host geometry and the device hierarchy in the
// cycle precision, plus the FP64 level-0 operator of the refinement
struct Medium {
  std::vector<std::vector<std::array<HostAxis, 3>>> axes;  // [pattern][level][axis]
  Hier<double> h64;
  Hier<float> h32;
  DevLevel<double> l0_64;
};

// one right-hand side
struct Slot {
  int med = 0;                       // its grid and medium
  double omu = 0;
  std::vector<void*> u, f, r;        // per level, working precision
  void* u64 = nullptr;               // level-0 FP64 state (aliases u/f/r[0] in FP64 mode)
  void* f64 = nullptr;
  void* r64 = nullptr;               // FP64 mode: r[0]; FP32-storage cycles: blend_guess scratch (lazy)
  std::vector<LineFactors> facs;     // [(pattern*lmax + level)*3 + axis]; level 0 under pattern 0
  void* lscr = nullptr;              // right-hand side / substitution records of the line solves
  double* partial = nullptr;         // residual-norm partial sums
  double rnorm0 = 0;
  SolveStats* st = nullptr;
  double seen_omu[3] = {-1.0, -1.0, -1.0};  // frequency whose factors a cycle of pattern p has built
};

class Solver {
 public:
  static constexpr int MAX_SLOTS = 64;   // active slots are tracked in a 64-bit mask
  // x1..x3: float node coordinates (n+1); sig: float [n3][n2][n1] conductivities
  // (host pointers when sig_on_device == 0, device pointers otherwise);
  // omegas: the angular frequency of each of the nslots right-hand sides
  Solver(int n1, int n2, int n3, const float* x1, const float* x2, const float* x3, const float* sig11,
         const float* sig22, const float* sig33, int sig_on_device, const float* omegas, int nslots,
         const SolveParams& p, hipStream_t stream);
  // nmed grids and media of one multigrid size (e.g. the shots of a survey,
  // each regridded about its own source); slot k solves on grid slot_med[k]
  // at omegas[k].  Slots on different grids still share every launch.
  Solver(int n1, int n2, int n3, int nmed, const float* const* x1, const float* const* x2, const float* const* x3,
         const float* const* sig11, const float* const* sig22, const float* const* sig33, int sig_on_device,
         const int* slot_med, const float* omegas, int nslots, const SolveParams& p, hipStream_t stream);
  ~Solver();
  Solver(const Solver&) = delete;
  Solver& operator=(const Solver&) = delete;

  // b, x, h: device vectors of 3(n1+1)(n2+1)(n3+1) complex doubles.  b is the
  // source term (before the dual-volume scaling); x receives E, h receives H
  // (h may alias b).  x0 (device, may be null) is an initial guess.
  int solve(const void* b, void* x, void* h, const void* x0, SolveStats* st);
  // slots 0..n-1 together; x0 may be null or hold null entries
  int solve_batch(int n, const void* const* b, void* const* x, void* const* h, const void* const* x0,
                  SolveStats* st);

  long nvec() const { return 3 * nc0_; }
  int lmax() const { return lmax_; }
  int npatterns() const { return npat_; }
  int nslots() const { return (int)slots_.size(); }
  double omu(int slot = 0) const { return slots_[slot].omu; }
  // the hierarchy depends on the grid and medium only: switch frequency
  void set_omega(float omega) { set_omega(0, omega); }
  void set_omega(int slot, float omega);
  // cycle limit of the following solves
  void set_ncycle(int ncycle) { p_.ncycle = ncycle; }
  // cycle k of the following solves uses semi-coarsening pattern
  // (first + k) % 3 (default 0, the reference's order)
  void set_first_cycle(int first) { first_cycle_ = first; }
  const int* dims(int pat, int lev) const { return dims_[pat][lev].data(); }

  // kernel-level access for the element-wise checks (hierarchies of slot 0's grid)
  const Hier<double>& h64() const { return meds_[slots_[0].med].h64; }
  const Hier<float>& h32() const { return meds_[slots_[0].med].h32; }
  int nmedia() const { return (int)meds_.size(); }
  hipStream_t stream() const { return s_; }
  // one line sweep (all four colours) along `axis` on vectors u, f of level
  // `lev` of pattern `pat`, with the frequency of slot 0, in the working
  // precision T of the solver
  template <typename T> void line_sweep_on(int pat, int lev, int axis, int iter, void* u, const void* f);
  // average time of one launch (ms) and compulsory bytes of the level-0
  // kernels of slot 0 in the cycle precision, each launched reps times:
  // residual, restriction and prolongation (pattern 0), a full line sweep
  // (four colours) along x, y, z; returns 6
  int kernel_rates(int reps, double* ms, double* bytes);
  // Autotuning of the line sweeps: for every launch shape a solve of all
  // slots issues (levels, axes, colours, packs of up to LINE_PACK_MAX
  // slots) and that the process-wide table does not hold yet, times each
  // substitution choice (one-lane, quad-lane fused, quad-lane with separate
  // right-hand sides) reps times on the slots' own level vectors and records
  // the fastest.  The vectors' contents are overwritten (the next solve sets
  // them again); captured cycle graphs are dropped.  Returns the number of
  // shapes tuned.
  int autotune(int reps);
  // Blends guesses with multigrid iterates before a solve from them: for
  // slots 0..n-1 (b: source terms as for solve_batch, x1: an iterate, xg: a
  // guess, overwritten) xg <- x1 + beta (xg - x1) with the complex beta that
  // minimises ||f - A xg|| (f = b scaled by the dual volumes), so the blend's
  // residual is never above x1's; beta = 0 when xg - x1 leaves the residual
  // unchanged or a norm is not finite.  beta: 2n doubles (re, im).
  void blend_guess(int n, const void* const* b, const void* const* x1, void* const* xg, double* beta);

 private:
  template <typename T> int autotune_t(int reps);
  void init(int n1, int n2, int n3, int nmed, const float* const* x1, const float* const* x2,
            const float* const* x3, const float* const* sig11, const float* const* sig22,
            const float* const* sig33, int sig_on_device, const int* slot_med, const float* omegas, int nslots);
  template <typename T> Hier<T>& hier(int med);
  template <typename T> void build_hier(Hier<T>& H, const std::vector<std::vector<double*>>& cells, int nlev,
                                        int med);
  template <typename T> void alloc_vectors(Slot& S, const Hier<T>& H);
  template <typename T> void smooth(int pat, int lev, int iter);
  template <typename T> void line_sweep(int pat, int lev, int axis, int iter);
  template <typename T> void point_sweep(int pat, int lev, int iter);
  template <typename T> void kernel_rates_t(int reps, double* ms, double* bytes);
  template <typename T> bool coarsens(int pat, int lev);
  bool exact_coarsest(int pat) const;
  template <typename T> LineFactors& factors(Slot& S, int pat, int lev, int axis);
  template <typename T> bool v_cycle(int pat, int lev, bool norm_here, bool zero_guess);
  template <typename T> bool f_cycle(int pat, int lev, bool norm_here, bool zero_guess);
  template <typename T> bool w_cycle(int pat, int lev, bool norm_here, bool zero_guess);
  template <typename T> bool coarse_correction(int pat, int lev, bool norm_here, bool zero_guess, int kind);
  // level-0 FP64 residual norms of the active slots (see norms); retires
  // the converged ones and returns true when none is left
  bool check_active(int pat);
  // one cycle after its convergence test, as a replayed graph when possible
  void cycle_body(int pat);
  void cycle_body_eager(int pat);
  void drop_graphs();
  // ||f64 - A u64|| of the slots `which`, batched.  FP64 cycles leave the
  // residual in r64 (= r[0]).  FP32-storage cycles leave it demoted in f[0]
  // (the next cycle's right-hand side); with `add` each slot's pending FP32
  // correction u[0] is added to u64 first.
  void norms(const std::vector<int>& which, bool add, std::vector<double>& out);

  int n_[3];
  long nc0_;
  int lmax_, npat_;
  SolveParams p_;
  hipStream_t s_;
  std::vector<std::vector<std::array<int, 3>>> dims_, scs_;   // [pattern][level]
  std::vector<Medium> meds_;
  std::vector<Slot> slots_;
  size_t lscr_bytes_ = 0;
  double* dnorm_ = nullptr;   // one per slot
  double* part_all_ = nullptr;   // residual-norm partials of MG_BATCH_MAX slots
  double* blend_part_ = nullptr;   // blend_guess: partial sums and their total (allocated on first use)
  double* blend_dots_ = nullptr;
  double* hnorm_ = nullptr;   // pinned host, one per slot
  std::vector<void*> owned_;
  // state of the running batch
  std::vector<int> act_;      // active slots
  int icycle_ = 0;
  int first_cycle_ = 0;
  bool r0_ready_ = false;     // r[0] holds f - A u of the level-0 test (FP64 form)
  bool rnorm0_in_cycle_ = false;  // rnorm0 is the cycle-0 in-cycle norm (FP64 V/W cycles)
  bool pending_ = false;      // FP32 cycles: the active slots' u[0] holds a correction not yet in u64
  // captured cycle bodies, keyed by pattern and active-slot mask; valid for
  // the slots' frequencies at capture time
  std::map<std::pair<int, unsigned long long>, hipGraphExec_t> graphs_;
  std::vector<double> graph_omu_;
};

}  // namespace emmi
