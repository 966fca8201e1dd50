// mg_kernels.h -- launchers of the gfx950 multigrid kernels (mg_kernels.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <cstddef>
#include "emmi_dev.h"

namespace emmi {

template <typename T>
struct Transfer {
  int sc[3];
  const double* wr[3];
};

// The per-slot vectors of one launch over several right-hand sides whose
// levels have the same shape (grid y = slot): in0/in1 are read, out0/out1
// written, omu is the slot's omega mu0.  Each slot may have its own grid and
// medium (the shots of a survey are regridded about their own sources): geo
// is the base of the slot's geometry block (W, IW, H, X, D per axis, gstride
// apart), sbar its edge medium and wr the base of its transfer weights
// (wstride apart per axis).
constexpr int MG_BATCH_MAX = 16;
template <typename T>
struct SlotVecs {
  const typename cvec<T>::type* in0[MG_BATCH_MAX];
  const typename cvec<T>::type* in1[MG_BATCH_MAX];
  typename cvec<T>::type* out0[MG_BATCH_MAX];
  typename cvec<T>::type* out1[MG_BATCH_MAX];
  double omu[MG_BATCH_MAX];
  const double* geo[MG_BATCH_MAX];
  const T* sbar[MG_BATCH_MAX];
  const double* wr[MG_BATCH_MAX];
  long gstride, wstride;
};
static_assert(sizeof(SlotVecs<double>) + 2 * sizeof(Level<double>) + sizeof(Transfer<double>) <= 4096,
              "batched launch arguments exceed 4 KiB");

// Vectors of the convergence test of an FP64 refinement step (k_refine_test):
// per slot the FP64 iterate u, the FP64 right-hand side f and the FP32
// residual r32.
struct RefineVecs {
  const double2* u[MG_BATCH_MAX];
  const double2* f[MG_BATCH_MAX];
  float2* r32[MG_BATCH_MAX];
  double omu[MG_BATCH_MAX];
  const double* geo[MG_BATCH_MAX];
  const double* sbar[MG_BATCH_MAX];
  long gstride;
};
static_assert(sizeof(RefineVecs) + sizeof(Level<double>) + sizeof(double*) <= 4096,
              "refinement-test launch arguments exceed 4 KiB");

template <typename T>
void launch_residual(const Level<T>& L, const void* u, const void* f, void* r, double* partial, hipStream_t s);
// For n slots of level-0 shape (Ls[k]: slot k's FP64 level 0): r = f_k - A u_k
// written demoted to FP32 into r32_k, and the |r|^2 partials of slot k at
// partial + k residual_blocks<double>(n).  One launch.
void launch_refine_test(const Level<double>* Ls, int n, const void* const* u, const void* const* f, void* const* r32,
                        double* partial, hipStream_t s);
// (partial, when non-null: slot k's |r|^2 partials at partial + k residual_blocks<T>(n))
template <typename T>
void launch_residual_batch(const Level<T>* Ls, int n, const void* const* u, const void* const* f, void* const* r,
                           const double* omu, hipStream_t s, double* partial = nullptr);
// one vertex-patch Gauss-Seidel sweep (src/gmg.c:116-336) of u_k against f_k
// for n slots of levels of one shape: 8 node colours, reversed for odd iter
template <typename T>
void launch_point_gs_batch(const Level<T>* Ls, int n, void* const* u, const void* const* f, const double* omu,
                           int iter, hipStream_t s);
// fc_k = R rf_k, uc_k zeroed, for n slots, one launch (F, C: the shapes;
// Xs[k]: slot k's transfer weights)
template <typename T>
void launch_restrict_batch(const Level<T>& F, const Level<T>& C, const Transfer<T>* Xs, int n, const void* const* rf,
                           void* const* fc, void* const* uc, hipStream_t s);
// uf_k += P uc_k for n slots, one launch
template <typename T>
void launch_prolong_batch(const Level<T>& F, const Level<T>& C, const Transfer<T>* Xs, int n, const void* const* uc,
                          void* const* uf, hipStream_t s);
// partial sums per slot of a residual launch over a level of shape n with
// storage T, and an upper bound for either storage
template <typename T>
int residual_blocks(const int* n);
int residual_blocks_max(const int* n);
void launch_sum_partials(const double* partial, int n, double* out, hipStream_t s);
// out[k] = the sum of the n partials at partial + k n, k < nslots (one launch)
void launch_sum_partials_batch(const double* partial, int n, int nslots, double* out, hipStream_t s);
template <typename T>
// fc = R rf; uc (may be null) is zeroed on the coarse level in the same pass
void launch_restrict(const Level<T>& F, const Level<T>& C, const Transfer<T>& X, const void* rf, void* fc, void* uc,
                     hipStream_t s);
template <typename T>
void launch_prolong(const Level<T>& F, const Level<T>& C, const Transfer<T>& X, const void* uc, void* uf, hipStream_t s);
template <typename T>
void launch_zero(void* v, long n, hipStream_t s);
template <typename T>
void launch_scale_rhs(const Level<T>& L, const double* D0, const double* D1, const double* D2, const void* b, void* f, hipStream_t s);
template <typename T>
void launch_curl_h(const Level<T>& L, const void* u, void* h, hipStream_t s);
// out3 = (re, im of d^H r1, |d|^2) with d = r1 - rg over n complex doubles;
// part holds 3 * blend_blocks(n) doubles
void launch_blend_dots(const void* r1, const void* rg, long n, double* part, double* out3, hipStream_t s);
int blend_blocks(long n);
// xg = x1 + beta (xg - x1) over n complex doubles
void launch_blend(void* xg, const void* x1, double br, double bi, long n, hipStream_t s);
void launch_promote_add(void* hi, const void* lo, long n, hipStream_t s);
void launch_sigma_vol0(int n1, int n2, int n3, const double* W0, const double* W1, const double* W2, const float* sig,
                       double* out, hipStream_t s);
void launch_coarsen_cells(int fn1, int fn2, int n1, int n2, int n3, const int* sc, const double* fine, double* out,
                          hipStream_t s);
template <typename T>
void launch_edge_sigma(int n1, int n2, int n3, const double* s11, const double* s22, const double* s33, T* sbar,
                       hipStream_t s);

}  // namespace emmi
