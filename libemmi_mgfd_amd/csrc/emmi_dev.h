// emmi_dev.h -- device-side types and helpers shared by the HIP kernels.
//
// Field vectors keep the reference's layout (src/gmg.c:1780-1782):
// [3][n3+1][n2+1][n1+1] complex, i fastest, component-major.
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

namespace emmi {

template <typename T> struct cvec;
template <> struct cvec<double> { using type = double2; };
template <> struct cvec<float> { using type = float2; };

// complex value with explicit real arithmetic (no library complex type)
template <typename T>
struct cx {
  T re, im;
  __host__ __device__ cx() : re(0), im(0) {}
  __host__ __device__ cx(T r, T i = 0) : re(r), im(i) {}
};

template <typename T> __host__ __device__ __forceinline__ cx<T> operator+(cx<T> a, cx<T> b) { return cx<T>(a.re + b.re, a.im + b.im); }
template <typename T> __host__ __device__ __forceinline__ cx<T> operator-(cx<T> a, cx<T> b) { return cx<T>(a.re - b.re, a.im - b.im); }
template <typename T> __host__ __device__ __forceinline__ cx<T> operator-(cx<T> a) { return cx<T>(-a.re, -a.im); }
template <typename T> __host__ __device__ __forceinline__ cx<T> operator*(cx<T> a, cx<T> b) {
  return cx<T>(a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re);
}
template <typename T> __host__ __device__ __forceinline__ cx<T> operator*(cx<T> a, T s) { return cx<T>(a.re * s, a.im * s); }
template <typename T> __host__ __device__ __forceinline__ cx<T> operator*(T s, cx<T> a) { return cx<T>(a.re * s, a.im * s); }
template <typename T> __host__ __device__ __forceinline__ cx<T>& operator+=(cx<T>& a, cx<T> b) { a.re += b.re; a.im += b.im; return a; }
template <typename T> __host__ __device__ __forceinline__ cx<T>& operator-=(cx<T>& a, cx<T> b) { a.re -= b.re; a.im -= b.im; return a; }
// 1/a
template <typename T> __host__ __device__ __forceinline__ cx<T> cinv(cx<T> a) {
  T s = T(1) / (a.re * a.re + a.im * a.im);
  return cx<T>(a.re * s, -a.im * s);
}
// a * (-i s): the mass term of the curl-curl rows
template <typename T> __host__ __device__ __forceinline__ cx<T> mul_minus_i(cx<T> a, T s) { return cx<T>(a.im * s, -a.re * s); }

using cxd = cx<double>;

// loads widen to double and stores narrow from double: every kernel computes
// in FP64 whatever the storage precision of its vectors
template <typename T>
__device__ __forceinline__ cxd ld(const typename cvec<T>::type* p, long i) {
  typename cvec<T>::type v = p[i];
  return cxd((double)v.x, (double)v.y);
}
template <typename T>
__device__ __forceinline__ void st(typename cvec<T>::type* p, long i, cxd v) {
  typename cvec<T>::type w;
  w.x = (T)v.re; w.y = (T)v.im;
  p[i] = w;
}

// Workgroups are dispatched round-robin over the 8 XCDs of the chip (block b
// runs on XCD b % 8), each XCD with its own L2.  xcd_block maps the dispatch
// index b of nb blocks to a logical block so that XCD x takes the x-th
// contiguous eighth of the logical blocks: a stencil's neighbours along the
// slowest axis are then loaded by the same XCD and found in its L2.  It is a
// bijection of [0, nb).
__device__ __forceinline__ long xcd_block(long b, long nb) {
  const long x = b & 7, q = b >> 3;
  const long base = nb >> 3, extra = nb & 7;
  return x * base + (x < extra ? x : extra) + q;
}

// (i, j, k) of the flat index t of a [n3+1][n2+1][n1+1] array with row
// length es1 = n1+1 and n2p1 = n2+1 rows per plane, in 32-bit arithmetic
// (level arrays hold fewer than 2^32 entries; 64-bit division is emulated
// with long instruction sequences).
__device__ __forceinline__ void ijk_of(long t, long es1, int n2p1, int& i, int& j, int& k) {
  const unsigned tu = (unsigned)t, e1 = (unsigned)es1, m = (unsigned)n2p1;
  const unsigned q = tu / e1;
  i = (int)(tu - q * e1);
  k = (int)(q / m);
  j = (int)(q - (unsigned)k * m);
}

// Geometry of one multigrid level as the kernels see it.  The 1D geometry is
// FP64 whatever the storage precision T of the vectors and of sbar: the
// discrete curl of a discrete gradient cancels exactly only when every
// operator of the hierarchy uses the same widths, so FP32-storage cycles see
// the FP64 operator's widths, not rounded ones.
//   W[a][i]  cell width along axis a (i in [0, n_a])
//   IW[a][i] 1/W[a][i]
//   H[a][i]  dual width 0.5*(W[a][i] + W[a][i-1]) (i in [1, n_a])
//   X[a][i]  node coordinate
//   sbar     per-edge mean of sigma*volume over the 4 cells around the edge,
//            [3][n3+1][n2+1][n1+1], zero off interior edges
template <typename T>
struct Level {
  int n[3];
  long es[3];
  long nc;
  const double* W[3];
  const double* IW[3];
  const double* H[3];
  const double* X[3];
  const T* sbar;
  double omu;
};

__host__ __device__ __forceinline__ int ax_b(int a) { return a == 2 ? 0 : a + 1; }
__host__ __device__ __forceinline__ int ax_c(int a) { return a == 0 ? 2 : a - 1; }

}  // namespace emmi
