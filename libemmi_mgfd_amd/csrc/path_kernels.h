// path_kernels.h -- launchers of the kernels around the solve: source scatter,
// receiver gathers, E on the model grid and the gradient.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>

namespace emmi {
// Static device scratch of the entry point that allocates nothing
// (emmi_inject_adj_source_dev_noalloc): a file-scope __device__ array of
// NOALLOC_SCRATCH_BYTES in every GPU's memory for the life of the process.
// noalloc_scratch() returns its address on the current device (looked up
// once per device with hipGetSymbolAddress); throws on failure.
constexpr size_t NOALLOC_SCRATCH_BYTES = (size_t)2 << 20;
void* noalloc_scratch();

void launch_zero_c128(void* v, long n, hipStream_t s);
void launch_scatter_set(void* v, const long* idx, const void* val, int n, hipStream_t s);
void launch_gather(const void* v, const long* idx, void* out, int n, hipStream_t s);

// per-axis cell-centre interpolation tables (device pointers), as
// CentreTables holds them on the host
struct CentreDev {
  const int *i1, *i1s, *i2, *i2s, *i3, *i3s;
  const float *w1, *w1s, *w2, *w2s, *w3, *w3s;
  int nx, ny, nz;
};
// E1, E2, E3 of one frequency on the model's cell centres: e = [3][nz][ny][nx]
// complex float, from the field vector x (complex double) of a grid n1 x n2 x n3;
// s33: MG-grid sigma33 [n3][n2][n1]; r33: model rho33 [nz][ny][nx]
void launch_efield(const CentreDev& T, const void* x, int n1, int n2, int n3, const float* s33, const float* r33,
                   void* e, hipStream_t s);

// H1, H2, H3 of one frequency on the model's cell centres: e = [3][nz][ny][nx]
// complex float, from the magnetic field vector h (complex double)
void launch_hfield(const CentreDev& T, const void* h, int n1, int n2, int n3, void* e, hipStream_t s);

// b += curl M over the nodes 1..n of each axis (src/inject_extract.c:303-312);
// m = the dipole vector, d1..d3 the dual widths (n+1 floats each)
void launch_add_curl_m(void* b, const void* m, int n1, int n2, int n3, const float* d1, const float* d2,
                       const float* d3, hipStream_t s);

constexpr int GRAD_MAX_FREQ = 32;
constexpr int GRAD_MAX_PAR = 8;
struct GradArgs {
  const void* ef[GRAD_MAX_FREQ];   // forward E per frequency, [3][nz][ny][nx] complex float
  const void* ea[GRAD_MAX_FREQ];   // adjoint E per frequency
  int nfreq, npar, preco;
  int idxpar[GRAD_MAX_PAR];
  int nx, ny, nz;
  float s, dzf;                    // dx*dy*dz; dz
  double oz, dz;
  const float *rho11, *rho33, *bathy;
  float *grad, *phess;             // [npar][nz][ny][nx]
};
void launch_gradient(const GradArgs& a, hipStream_t s);
// zero grad[*][box] for each box (i1min, i1max, i2min, i2max, i3min, i3max)
void launch_mute(float* grad, int npar, int nx, int ny, int nz, const int* boxes, int nbox, int maxvol, hipStream_t s);
}  // namespace emmi
