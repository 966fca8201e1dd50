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
// a x - b y per component with one rounded product: fma(a, x, -(b y))
__device__ __forceinline__ cxd fms2(const cxd& a, double x, const cxd& b, double y) {
  return cxd(fma(a.re, x, -(b.re * y)), fma(a.im, x, -(b.im * y)));
}

// get(x) returns the value at flat index x of the edge vector (component
// blocks L.nc apart); o is the flat index of node p.  Every product is an
// explicit fma or stands alone, so the compiler has nothing left to contract
// and each kernel that calls this (whatever its loads) rounds alike.
template <typename T, int A, typename Get>
__device__ __forceinline__ cxd apply_row_get(const Level<T>& L, const Get& get, int pa, int pb, int pc, long o) {
  constexpr int B = (A + 1) % 3, C = (A + 2) % 3;
  const long sa = L.es[A], sb = L.es[B], sc = L.es[C];
  const long oa = A * L.nc + o, ob = B * L.nc + o, oc = C * L.nc + o;
  const double iWa = L.IW[A][pa], iWb = L.IW[B][pb], iWbm = L.IW[B][pb - 1];
  const double iWc = L.IW[C][pc], iWcm = L.IW[C][pc - 1];
  const cxd ea = get(oa);
  const cxd c1 = fms2(get(ob + sa) - get(ob), iWa, get(oa + sb) - ea, iWb);
  const cxd c2 = fms2(get(ob + sa - sb) - get(ob - sb), iWa, ea - get(oa - sb), iWbm);
  const cxd c3 = fms2(get(oa + sc) - ea, iWc, get(oc + sa) - get(oc), iWa);
  const cxd c4 = fms2(ea - get(oa - sc), iWcm, get(oc + sa - sc) - get(oc - sc), iWa);
  const cxd q = fms2(c1 - c2, (double)L.H[C][pc], c3 - c4, (double)L.H[B][pb]);
  // W_a q - i omu sbar ea
  const double wa = L.W[A][pa], s = L.omu * (double)L.sbar[oa];
  return cxd(fma(q.re, wa, ea.im * s), fma(q.im, wa, -(ea.re * s)));
}

template <typename T, int A>
__device__ __forceinline__ cxd apply_row(const Level<T>& L, const typename cvec<T>::type* __restrict__ u,
                                           int pa, int pb, int pc, long o) {
  return apply_row_get<T, A>(L, [u](long x) { return ld<T>(u, x); }, pa, pb, pc, o);
}

template <int A>
__device__ __forceinline__ bool interior(const int* n, int i, int j, int k) {
  const int p[3] = {i, j, k};
  constexpr int B = (A + 1) % 3, C = (A + 2) % 3;
  return p[A] < n[A] && p[B] >= 1 && p[B] <= n[B] - 1 && p[C] >= 1 && p[C] <= n[C] - 1;
}

}  // namespace emmi
