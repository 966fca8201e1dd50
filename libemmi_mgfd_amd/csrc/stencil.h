// stencil.h -- one row of the curl-curl operator, shared by the kernels.
//
//   (A u)_a(p) = W_a(p_a) * [ H_c(p_c) (C_c(p) - C_c(p - e_b))
//                           - H_b(p_b) (C_b(p) - C_b(p - e_c)) ]
//                - i omega mu0 sbar_a(p) E_a(p)
//   C_c(q) = (E_b(q+e_a) - E_b(q))/W_a - (E_a(q+e_b) - E_a(q))/W_b
//   C_b(q) = (E_a(q+e_c) - E_a(q))/W_c - (E_c(q+e_a) - E_c(q))/W_a
// with (a, b, c) cyclic; this is src/gmg.c:1365-1405 with mu_r = 1 and the
// 2-cell means of invmur*vol written as products of 1D widths.
#pragma once

#include "emmi_dev.h"

namespace emmi {

// ---------------------------------------------------------------------------
// operator row
// ---------------------------------------------------------------------------
template <typename T, int A>
__device__ __forceinline__ cxd apply_row(const Level<T>& L, const typename cvec<T>::type* __restrict__ u,
                                           int pa, int pb, int pc, long o) {
  constexpr int B = (A + 1) % 3, C = (A + 2) % 3;
  const long sa = L.es[A], sb = L.es[B], sc = L.es[C];
  const typename cvec<T>::type* ua = u + A * L.nc + o;
  const typename cvec<T>::type* ub = u + B * L.nc + o;
  const typename cvec<T>::type* uc = u + C * L.nc + o;
  const double iWa = L.IW[A][pa], iWb = L.IW[B][pb], iWbm = L.IW[B][pb - 1];
  const double iWc = L.IW[C][pc], iWcm = L.IW[C][pc - 1];
  const cxd ea = ld<T>(ua, 0);
  const cxd c1 = (ld<T>(ub, sa) - ld<T>(ub, 0)) * iWa - (ld<T>(ua, sb) - ea) * iWb;
  const cxd c2 = (ld<T>(ub, sa - sb) - ld<T>(ub, -sb)) * iWa - (ea - ld<T>(ua, -sb)) * iWbm;
  const cxd c3 = (ld<T>(ua, sc) - ea) * iWc - (ld<T>(uc, sa) - ld<T>(uc, 0)) * iWa;
  const cxd c4 = (ea - ld<T>(ua, -sc)) * iWcm - (ld<T>(uc, sa - sc) - ld<T>(uc, -sc)) * iWa;
  cxd r = ((c1 - c2) * (double)L.H[C][pc] - (c3 - c4) * (double)L.H[B][pb]) * (double)L.W[A][pa];
  return r + mul_minus_i(ea, L.omu * (double)L.sbar[A * L.nc + o]);
}

template <int A>
__device__ __forceinline__ bool interior(const int* n, int i, int j, int k) {
  const int p[3] = {i, j, k};
  constexpr int B = (A + 1) % 3, C = (A + 2) % 3;
  return p[A] < n[A] && p[B] >= 1 && p[B] <= n[B] - 1 && p[C] >= 1 && p[C] <= n[C] - 1;
}

}  // namespace emmi
