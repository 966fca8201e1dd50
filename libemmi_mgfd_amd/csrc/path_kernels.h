// path_kernels.h -- launchers of the source-scatter and receiver-gather kernels.
#pragma once

#include <hip/hip_runtime.h>

namespace emmi {
void launch_zero_c128(void* v, long n, hipStream_t s);
void launch_scatter_set(void* v, const long* idx, const void* val, int n, hipStream_t s);
void launch_gather(const void* v, const long* idx, void* out, int n, hipStream_t s);
}  // namespace emmi
