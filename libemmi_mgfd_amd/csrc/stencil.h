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

// The row's arithmetic on its thirteen values
//   v = E_a(o), E_b(o+e_a), E_b(o), E_a(o+e_b), E_b(o+e_a-e_b), E_b(o-e_b),
//       E_a(o-e_b), E_a(o+e_c), E_c(o+e_a), E_c(o), E_a(o-e_c),
//       E_c(o+e_a-e_c), E_c(o-e_c)
// and its geometry.  Every product is an explicit fma or stands alone, so
// the compiler has nothing left to contract and each kernel that calls this
// (whatever its loads) rounds alike.
struct RowGeo {
  double iWa, iWb, iWbm, iWc, iWcm;   // 1/W_a(p_a), 1/W_b(p_b), 1/W_b(p_b-1), 1/W_c(p_c), 1/W_c(p_c-1)
  double Hc, Hb, wa, s;                // H_c(p_c), H_b(p_b), W_a(p_a), omega mu0 sbar_a(p)
};
__device__ __forceinline__ cxd row_arith(const cxd (&v)[13], const RowGeo& g) {
  const cxd ea = v[0];
  const cxd c1 = fms2(v[1] - v[2], g.iWa, v[3] - ea, g.iWb);
  const cxd c2 = fms2(v[4] - v[5], g.iWa, ea - v[6], g.iWbm);
  const cxd c3 = fms2(v[7] - ea, g.iWc, v[8] - v[9], g.iWa);
  const cxd c4 = fms2(ea - v[10], g.iWcm, v[11] - v[12], g.iWa);
  const cxd q = fms2(c1 - c2, g.Hc, c3 - c4, g.Hb);
  // W_a q - i omu sbar ea
  return cxd(fma(q.re, g.wa, ea.im * g.s), fma(q.im, g.wa, -(ea.re * g.s)));
}

// get(x) returns the value at flat index x of the edge vector (component
// blocks L.nc apart); o is the flat index of node p.
template <typename T, int A, typename Get>
__device__ __forceinline__ cxd apply_row_get(const Level<T>& L, const Get& get, int pa, int pb, int pc, long o) {
  constexpr int B = (A + 1) % 3, C = (A + 2) % 3;
  const long sa = L.es[A], sb = L.es[B], sc = L.es[C];
  const long oa = A * L.nc + o, ob = B * L.nc + o, oc = C * L.nc + o;
  const RowGeo g{L.IW[A][pa], L.IW[B][pb], L.IW[B][pb - 1], L.IW[C][pc], L.IW[C][pc - 1],
                 L.H[C][pc],  L.H[B][pb],  L.W[A][pa],      L.omu * (double)L.sbar[oa]};
  const cxd v[13] = {get(oa),           get(ob + sa), get(ob),      get(oa + sb), get(ob + sa - sb),
                     get(ob - sb),      get(oa - sb), get(oa + sc), get(oc + sa), get(oc),
                     get(oa - sc),      get(oc + sa - sc), get(oc - sc)};
  return row_arith(v, g);
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
