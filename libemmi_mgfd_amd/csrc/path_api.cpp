// path_api.cpp -- device-resident source injection and receiver extraction.
//
// The arithmetic of both steps stays on the host in the reference's exact
// float/double order (host_path.cpp); only the 24 source entries per source
// and the 48 field entries per receiver cross PCIe, never the field vectors.
#include <hip/hip_runtime.h>

#include <stdexcept>
#include <string>
#include <vector>

#include "emmi_mgfd.h"
#include "host_path.h"
#include "path_kernels.h"

namespace {
thread_local std::string g_path_err;

struct DevScratch {
  void* p = nullptr;
  size_t cap = 0;
  void* get(size_t bytes) {
    if (bytes > cap) {
      if (p) hipFree(p);
      if (hipMalloc(&p, bytes) != hipSuccess) throw std::runtime_error("hipMalloc failed");
      cap = bytes;
    }
    return p;
  }
  ~DevScratch() { if (p) hipFree(p); }
};
thread_local DevScratch g_idx, g_val;

void check(hipError_t e, const char* w) {
  if (e != hipSuccess) throw std::runtime_error(std::string(w) + ": " + hipGetErrorString(e));
}
}  // namespace

extern "C" int emmi_inject_source_dev(acq_t* acq, emf_t* emf, int ifreq, void* b_dev, void* stream) {
  try {
    hipStream_t s = (hipStream_t)stream;
    std::vector<long> idx;
    std::vector<emmi_c128> val;
    emmi::source_entries(acq, emf, ifreq, idx, val);
    const long n = 3L * (emf->n1 + 1) * (emf->n2 + 1) * (emf->n3 + 1);
    for (long i : idx)
      if (i < 0 || i >= n) throw std::runtime_error("source entry outside the vector");
    long* di = (long*)g_idx.get(sizeof(long) * (idx.size() + 1));
    void* dv = g_val.get(sizeof(emmi_c128) * (val.size() + 1));
    check(hipMemcpyAsync(di, idx.data(), sizeof(long) * idx.size(), hipMemcpyHostToDevice, s), "h2d");
    check(hipMemcpyAsync(dv, val.data(), sizeof(emmi_c128) * val.size(), hipMemcpyHostToDevice, s), "h2d");
    emmi::launch_zero_c128(b_dev, n, s);
    emmi::launch_scatter_set(b_dev, di, dv, (int)idx.size(), s);
    check(hipStreamSynchronize(s), "sync");
    return 0;
  } catch (const std::exception& e) {
    g_path_err = e.what();
    return -1;
  }
}

extern "C" int emmi_extract_emf_data_dev(acq_t* acq, emf_t* emf, int ifreq, const void* x_dev, const void* h_dev,
                                         void* stream) {
  try {
    hipStream_t s = (hipStream_t)stream;
    std::vector<long> ix, ib;
    emmi::receiver_indices(acq, emf, ix, ib);
    const long n = 3L * (emf->n1 + 1) * (emf->n2 + 1) * (emf->n3 + 1);
    for (long i : ix) if (i < 0 || i >= n) throw std::runtime_error("receiver entry outside the vector");
    for (long i : ib) if (i < 0 || i >= n) throw std::runtime_error("receiver entry outside the vector");
    const size_t nx = ix.size(), nbv = ib.size();
    long* di = (long*)g_idx.get(sizeof(long) * (nx + nbv + 1));
    emmi_c128* dv = (emmi_c128*)g_val.get(sizeof(emmi_c128) * (nx + nbv + 1));
    std::vector<long> all(ix);
    all.insert(all.end(), ib.begin(), ib.end());
    check(hipMemcpyAsync(di, all.data(), sizeof(long) * all.size(), hipMemcpyHostToDevice, s), "h2d");
    emmi::launch_gather(x_dev, di, dv, (int)nx, s);
    emmi::launch_gather(h_dev, di + nx, dv + nx, (int)nbv, s);
    std::vector<emmi_c128> vx(nx), vb(nbv);
    check(hipMemcpyAsync(vx.data(), dv, sizeof(emmi_c128) * nx, hipMemcpyDeviceToHost, s), "d2h");
    check(hipMemcpyAsync(vb.data(), dv + nx, sizeof(emmi_c128) * nbv, hipMemcpyDeviceToHost, s), "d2h");
    check(hipStreamSynchronize(s), "sync");
    emmi::receiver_data_from(acq, emf, ifreq, vx, vb);
    return 0;
  } catch (const std::exception& e) {
    g_path_err = e.what();
    return -1;
  }
}

extern "C" const char* emmi_path_last_error(void) { return g_path_err.c_str(); }
