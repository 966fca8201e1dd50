// path_kernels.hip -- device side of the steps around the solve: sparse
// source scatter and receiver gathers.
#include <hip/hip_runtime.h>

#include "path_kernels.h"

namespace emmi {

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

static unsigned nb(long n) { return (unsigned)((n + 255) / 256); }

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
