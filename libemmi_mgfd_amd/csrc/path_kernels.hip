// path_kernels.hip -- device side of the steps around the solve: sparse
// source scatter, receiver gathers, E on the model grid and the gradient.
// The last two follow the host path's float/double evaluation order with
// floating-point contraction off, so they reproduce it bit for bit (the
// pseudo-Hessian's hypot aside).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <mutex>
#include <stdexcept>
#include <string>

#include "path_kernels.h"

namespace emmi {

// The scratch of the entry point that allocates nothing.  In host code the
// name `scratch` stands for the symbol, not for a device address: host code
// uses only the address hipGetSymbolAddress returns for the current device.
static __device__ __align__(256) unsigned char scratch[NOALLOC_SCRATCH_BYTES];

void* noalloc_scratch() {
  constexpr int MAXDEV = 64;
  static void* addr[MAXDEV] = {};
  static std::mutex mu;
  int dev = -1;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) throw std::runtime_error(std::string("hipGetDevice: ") + hipGetErrorString(e));
  if (dev < 0 || dev >= MAXDEV) throw std::runtime_error("device index beyond the scratch table");
  std::lock_guard<std::mutex> lk(mu);
  if (!addr[dev]) {
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      throw std::runtime_error(std::string("hipGetSymbolAddress of the static scratch: ") + hipGetErrorString(e));
    addr[dev] = p;
  }
  return addr[dev];
}

__global__ void k_zero_c128(double2* __restrict__ v, long n) {
  const long t = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (t < n) v[t] = make_double2(0.0, 0.0);
}

__global__ void k_scatter_set(double2* __restrict__ v, const long* __restrict__ idx, const double2* __restrict__ val,
                              int n) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) v[idx[t]] = val[t];
}

__global__ void k_gather(const double2* __restrict__ v, const long* __restrict__ idx, double2* __restrict__ out, int n) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) out[t] = v[idx[t]];
}

struct cd {
  double re, im;
};
__device__ __forceinline__ cd cadd(cd a, cd b) { return cd{a.re + b.re, a.im + b.im}; }
__device__ __forceinline__ cd cmulr(cd a, double r) { return cd{a.re * r, a.im * r}; }
__device__ __forceinline__ cd ldz(const double2* __restrict__ v, long i) { const double2 t = v[i]; return cd{t.x, t.y}; }

// trilinear gather of a staggered component (src/inject_extract.c:504-560)
__device__ __forceinline__ cd gather8(const double2* __restrict__ v, long o, long s1, long s2, float w1f, float w2f,
                                      float w3f) {
#pragma clang fp contract(off)
  const double w1 = w1f, w2 = w2f, w3 = w3f;
  cd s{0, 0};
  s = cadd(s, cmulr(cmulr(cmulr(ldz(v, o), 1. - w1), 1. - w2), 1. - w3));
  s = cadd(s, cmulr(cmulr(cmulr(ldz(v, o + 1), w1), 1. - w2), 1. - w3));
  s = cadd(s, cmulr(cmulr(cmulr(ldz(v, o + s1), 1. - w1), w2), 1. - w3));
  s = cadd(s, cmulr(cmulr(cmulr(ldz(v, o + 1 + s1), w1), w2), 1. - w3));
  s = cadd(s, cmulr(cmulr(cmulr(ldz(v, o + s2), 1. - w1), 1. - w2), w3));
  s = cadd(s, cmulr(cmulr(cmulr(ldz(v, o + 1 + s2), w1), 1. - w2), w3));
  s = cadd(s, cmulr(cmulr(cmulr(ldz(v, o + s1 + s2), 1. - w1), w2), w3));
  s = cadd(s, cmulr(cmulr(cmulr(ldz(v, o + 1 + s1 + s2), w1), w2), w3));
  return s;
}

// E on the model's cell centres (src/inject_extract.c:477-635): Ex, Ey by
// trilinear gathers; Ez as the vertical current (four-cell mean of sigma33
// around each Ez edge times Ez) gathered and times rho33
__global__ void k_efield(CentreDev T, const double2* __restrict__ x, long n123, int n1, int n2,
                         const float* __restrict__ s33, const float* __restrict__ r33, float2* __restrict__ e) {
#pragma clang fp contract(off)
  const long nxyz = (long)T.nx * T.ny * T.nz;
  const long o = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (o >= nxyz) return;
  const int i = (int)(o % T.nx), j = (int)((o / T.nx) % T.ny), k = (int)(o / ((long)T.nx * T.ny));
  const long s1 = n1 + 1, s2 = (long)(n1 + 1) * (n2 + 1);
  cd v = gather8(x, T.i1s[i] + s1 * T.i2[j] + s2 * T.i3[k], s1, s2, T.w1s[i], T.w2[j], T.w3[k]);
  e[o] = make_float2((float)v.re, (float)v.im);
  v = gather8(x, n123 + T.i1[i] + s1 * T.i2s[j] + s2 * T.i3[k], s1, s2, T.w1[i], T.w2s[j], T.w3[k]);
  e[nxyz + o] = make_float2((float)v.re, (float)v.im);
  const int kk = T.i3s[k], jj = T.i2[j], ii = T.i1[i];
  const double wz = T.w3s[k], wy = T.w2[j], wx = T.w1[i];
  cd s{0, 0};
  auto S33 = [&](int a, int b, int c) { return s33[a + (long)n1 * (b + (long)n2 * c)]; };
  auto term = [&](int a, int b2, int c, double f1, double f2, double f3) {
    const float sg = (float)(0.25 * (double)(S33(a, b2, c) + S33(a, b2 - 1, c) + S33(a - 1, b2, c) + S33(a - 1, b2 - 1, c)));
    const cd xv = ldz(x, 2 * n123 + a + s1 * b2 + s2 * c);
    s = cadd(s, cmulr(cmulr(cmulr(cd{(double)sg * xv.re, (double)sg * xv.im}, f1), f2), f3));
  };
  term(ii, jj, kk, 1. - wx, 1. - wy, 1. - wz);
  term(ii + 1, jj, kk, wx, 1. - wy, 1. - wz);
  term(ii, jj + 1, kk, 1. - wx, wy, 1. - wz);
  term(ii + 1, jj + 1, kk, wx, wy, 1. - wz);
  term(ii, jj, kk + 1, 1. - wx, 1. - wy, wz);
  term(ii + 1, jj, kk + 1, wx, 1. - wy, wz);
  term(ii, jj + 1, kk + 1, 1. - wx, wy, wz);
  term(ii + 1, jj + 1, kk + 1, wx, wy, wz);
  const cd ez = cmulr(s, (double)r33[o]);
  e[2 * nxyz + o] = make_float2((float)ez.re, (float)ez.im);
}

// H on the model's cell centres (src/inject_extract.c:638-785): trilinear
// gathers of the three staggered components of h
__global__ void k_hfield(CentreDev T, const double2* __restrict__ h, long n123, int n1, int n2,
                         float2* __restrict__ e) {
#pragma clang fp contract(off)
  const long nxyz = (long)T.nx * T.ny * T.nz;
  const long o = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (o >= nxyz) return;
  const int i = (int)(o % T.nx), j = (int)((o / T.nx) % T.ny), k = (int)(o / ((long)T.nx * T.ny));
  const long s1 = n1 + 1, s2 = (long)(n1 + 1) * (n2 + 1);
  cd v = gather8(h, T.i1[i] + s1 * T.i2s[j] + s2 * T.i3s[k], s1, s2, T.w1[i], T.w2s[j], T.w3s[k]);
  e[o] = make_float2((float)v.re, (float)v.im);
  v = gather8(h, n123 + T.i1s[i] + s1 * T.i2[j] + s2 * T.i3s[k], s1, s2, T.w1s[i], T.w2[j], T.w3s[k]);
  e[nxyz + o] = make_float2((float)v.re, (float)v.im);
  v = gather8(h, 2 * n123 + T.i1s[i] + s1 * T.i2s[j] + s2 * T.i3[k], s1, s2, T.w1s[i], T.w2s[j], T.w3[k]);
  e[2 * nxyz + o] = make_float2((float)v.re, (float)v.im);
}

// b += curl M with the dual widths (mu_r = 1): one thread per node, which
// owns the three entries of b at that node
__global__ void k_add_curl_m(double2* __restrict__ b, const double2* __restrict__ m, int n1, int n2, int n3,
                             const float* __restrict__ d1, const float* __restrict__ d2, const float* __restrict__ d3) {
#pragma clang fp contract(off)
  const long nn = (long)n1 * n2 * n3;
  const long t = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (t >= nn) return;
  const int i = 1 + (int)(t % n1), j = 1 + (int)((t / n1) % n2), k = 1 + (int)(t / ((long)n1 * n2));
  const long s1 = n1 + 1, s2 = (long)(n1 + 1) * (n2 + 1), n123 = s2 * (n3 + 1);
  const long o = i + s1 * j + s2 * k;
  const double w1 = d1[i], w2 = d2[j], w3 = d3[k];
  const cd mx = ldz(m, o), my = ldz(m, n123 + o), mz = ldz(m, 2 * n123 + o);
  auto sub = [](cd a, cd c) { return cd{a.re - c.re, a.im - c.im}; };
  auto divr = [](cd a, double r) { return cd{a.re / r, a.im / r}; };
  auto add = [&](long at, cd v) { const double2 h = b[at]; b[at] = make_double2(h.x + v.re, h.y + v.im); };
  add(o, sub(divr(sub(mz, ldz(m, 2 * n123 + o - s1)), w2), divr(sub(my, ldz(m, n123 + o - s2)), w3)));
  add(n123 + o, sub(divr(sub(mx, ldz(m, o - s2)), w3), divr(sub(mz, ldz(m, 2 * n123 + o - 1)), w1)));
  add(2 * n123 + o, sub(divr(sub(my, ldz(m, n123 + o - 1)), w1), divr(sub(mx, ldz(m, o - s1)), w2)));
}

// real part of a float complex product, in float
__device__ __forceinline__ float re_mul(float2 a, float2 b) {
#pragma clang fp contract(off)
  return a.x * b.x - a.y * b.y;
}

// gradient and pseudo-Hessian per (parameter, cell) (src/build_gradient.c:13-62)
__global__ void k_gradient(GradArgs a) {
#pragma clang fp contract(off)
  const long nxyz = (long)a.nx * a.ny * a.nz;
  const long t = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (t >= nxyz * a.npar) return;
  const int kp = (int)(t / nxyz);
  const long o = t - (long)kp * nxyz;
  const int i1 = (int)(o % a.nx), i2 = (int)((o / a.nx) % a.ny), i3 = (int)(o / ((long)a.nx * a.ny));
  const int par = a.idxpar[kp];
  float g = 0.f;
  for (int f = 0; f < a.nfreq; f++) {
    const float2* ef = (const float2*)a.ef[f];
    const float2* ea = (const float2*)a.ea[f];
    if (par == 1) {
      g = (float)((double)g + (double)re_mul(ef[o], ea[o]));
      g = (float)((double)g + (double)re_mul(ef[nxyz + o], ea[nxyz + o]));
    }
    if (par == 2) g = (float)((double)g + (double)re_mul(ef[2 * nxyz + o], ea[2 * nxyz + o]));
  }
  if (par == 1) g *= a.s / a.rho11[o];
  if (par == 2) g *= a.s / a.rho33[o];
  if (a.oz + ((double)i3 + 0.5) * a.dz <= (double)(a.bathy[i1 + (long)a.nx * i2] + a.dzf)) g = 0.f;
  a.grad[t] = g;
  if (a.preco == 1) {
    float h = 0.f;
    for (int f = 0; f < a.nfreq; f++) {
      const float2* ef = (const float2*)a.ef[f];
      auto ab = [](float2 z) { return hypot((double)z.x, (double)z.y); };
      if (par == 1) {
        h = (float)((double)h + ab(ef[o]) * ab(ef[o]));
        h = (float)((double)h + ab(ef[nxyz + o]) * ab(ef[nxyz + o]));
      }
      if (par == 2) h = (float)((double)h + ab(ef[2 * nxyz + o]) * ab(ef[2 * nxyz + o]));
    }
    if (par == 1) h *= a.s / a.rho11[o];
    if (par == 2) h *= a.s / a.rho33[o];
    a.phess[t] = h;
  }
}

// mutes around sources and receivers (src/build_gradient.c:64-111)
__global__ void k_mute(float* __restrict__ grad, int npar, int nx, int ny, int nz, const int* __restrict__ boxes,
                       int maxvol) {
  const int b = blockIdx.y;
  const int* bx = boxes + 6 * b;
  const int n1 = bx[1] - bx[0] + 1, n2 = bx[3] - bx[2] + 1, n3 = bx[5] - bx[4] + 1;
  const long nxyz = (long)nx * ny * nz;
  for (long c = blockIdx.x * (long)blockDim.x + threadIdx.x; c < (long)n1 * n2 * n3; c += (long)gridDim.x * blockDim.x) {
    const int a1 = bx[0] + (int)(c % n1), a2 = bx[2] + (int)((c / n1) % n2), a3 = bx[4] + (int)(c / ((long)n1 * n2));
    const long o = a1 + (long)nx * (a2 + (long)ny * a3);
    for (int k = 0; k < npar; k++) grad[k * nxyz + o] = 0.f;
  }
}

static unsigned nb(long n) { return (unsigned)((n + 255) / 256); }

void launch_efield(const CentreDev& T, const void* x, int n1, int n2, int n3, const float* s33, const float* r33,
                   void* e, hipStream_t s) {
  const long nxyz = (long)T.nx * T.ny * T.nz;
  const long n123 = (long)(n1 + 1) * (n2 + 1) * (n3 + 1);
  hipLaunchKernelGGL(k_efield, dim3(nb(nxyz)), dim3(256), 0, s, T, (const double2*)x, n123, n1, n2, s33, r33,
                     (float2*)e);
}

void launch_hfield(const CentreDev& T, const void* h, int n1, int n2, int n3, void* e, hipStream_t s) {
  const long nxyz = (long)T.nx * T.ny * T.nz;
  const long n123 = (long)(n1 + 1) * (n2 + 1) * (n3 + 1);
  hipLaunchKernelGGL(k_hfield, dim3(nb(nxyz)), dim3(256), 0, s, T, (const double2*)h, n123, n1, n2, (float2*)e);
}

void launch_add_curl_m(void* b, const void* m, int n1, int n2, int n3, const float* d1, const float* d2,
                       const float* d3, hipStream_t s) {
  const long nn = (long)n1 * n2 * n3;
  hipLaunchKernelGGL(k_add_curl_m, dim3(nb(nn)), dim3(256), 0, s, (double2*)b, (const double2*)m, n1, n2, n3, d1, d2, d3);
}

void launch_gradient(const GradArgs& a, hipStream_t s) {
  const long n = (long)a.nx * a.ny * a.nz * a.npar;
  hipLaunchKernelGGL(k_gradient, dim3(nb(n)), dim3(256), 0, s, a);
}

void launch_mute(float* grad, int npar, int nx, int ny, int nz, const int* boxes, int nbox, int maxvol, hipStream_t s) {
  if (nbox <= 0 || maxvol <= 0) return;
  const unsigned gx = (unsigned)std::min<long>(nb(maxvol), 64);
  hipLaunchKernelGGL(k_mute, dim3(gx, nbox), dim3(256), 0, s, grad, npar, nx, ny, nz, boxes, maxvol);
}

void launch_zero_c128(void* v, long n, hipStream_t s) {
  hipLaunchKernelGGL(k_zero_c128, dim3(nb(n)), dim3(256), 0, s, (double2*)v, n);
}
void launch_scatter_set(void* v, const long* idx, const void* val, int n, hipStream_t s) {
  if (n > 0) hipLaunchKernelGGL(k_scatter_set, dim3(nb(n)), dim3(256), 0, s, (double2*)v, idx, (const double2*)val, n);
}
void launch_gather(const void* v, const long* idx, void* out, int n, hipStream_t s) {
  if (n > 0) hipLaunchKernelGGL(k_gather, dim3(nb(n)), dim3(256), 0, s, (const double2*)v, idx, (double2*)out, n);
}

}  // namespace emmi
