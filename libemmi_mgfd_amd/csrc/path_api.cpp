// path_api.cpp -- device-resident steps around the solve.
//
// Source injection and receiver extraction keep their arithmetic on the host
// in the reference's exact float/double order (host_path.cpp); only the 24
// source entries per source and the 48 field entries per receiver cross
// PCIe, never the field vectors.  E on the model grid and the gradient run
// on the device (path_kernels.hip); their index tables and the medium arrays
// they read go up per call.
#include <hip/hip_runtime.h>

#include <algorithm>
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
thread_local DevScratch g_idx, g_val, g_tab, g_med, g_box;

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

namespace {
// src/inject_extract.c:124-314: the scatters accumulated on the host in the
// reference's order, b += curl M on the device.  The device tables (entry
// indices, values, dual widths) go to library-allocated scratch, or, with
// noalloc, to the static scratch of this device.
int adj_source(acq_t* acq, emf_t* emf, int ifreq, void* x_dev, void* b_dev, void* stream, bool noalloc) {
  try {
    hipStream_t s = (hipStream_t)stream;
    std::vector<long> bi, xi;
    std::vector<emmi_c128> bv, xv;
    emmi::adj_entries(acq, emf, ifreq, bi, bv, xi, xv);
    const int n1 = emf->n1, n2 = emf->n2, n3 = emf->n3;
    const long n = 3L * (n1 + 1) * (n2 + 1) * (n3 + 1);
    for (long i : bi) if (i < 0 || i >= n) throw std::runtime_error("adjoint source entry outside the vector");
    for (long i : xi) if (i < 0 || i >= n) throw std::runtime_error("adjoint source entry outside the vector");
    const size_t nb = bi.size(), nx = xi.size(), nw = (size_t)n1 + n2 + n3 + 3;
    auto pad = [](size_t b) { return (b + 255) & ~(size_t)255; };
    const size_t bidx = pad(sizeof(long) * (nb + nx + 1)), bval = pad(sizeof(emmi_c128) * (nb + nx + 1));
    const size_t bw = pad(sizeof(float) * nw);
    long* di;
    emmi_c128* dv;
    float* dw;
    if (noalloc) {
      if (bidx + bval + bw > emmi::NOALLOC_SCRATCH_BYTES)
        throw std::runtime_error("the adjoint source of " + std::to_string(acq->nrec) + " receivers needs " +
                                 std::to_string(bidx + bval + bw) + " bytes of device tables; the static scratch " +
                                 "holds " + std::to_string(emmi::NOALLOC_SCRATCH_BYTES) +
                                 " (use emmi_inject_adj_source_dev)");
      char* base = (char*)emmi::noalloc_scratch();
      di = (long*)base;
      dv = (emmi_c128*)(base + bidx);
      dw = (float*)(base + bidx + bval);
    } else {
      di = (long*)g_idx.get(bidx);
      dv = (emmi_c128*)g_val.get(bval);
      dw = (float*)g_tab.get(bw);
    }
    std::vector<long> idx(bi);
    idx.insert(idx.end(), xi.begin(), xi.end());
    std::vector<emmi_c128> val(bv);
    val.insert(val.end(), xv.begin(), xv.end());
    check(hipMemcpyAsync(di, idx.data(), sizeof(long) * idx.size(), hipMemcpyHostToDevice, s), "h2d");
    check(hipMemcpyAsync(dv, val.data(), sizeof(emmi_c128) * val.size(), hipMemcpyHostToDevice, s), "h2d");
    std::vector<float> dd;
    dd.insert(dd.end(), emf->d1, emf->d1 + n1 + 1);
    dd.insert(dd.end(), emf->d2, emf->d2 + n2 + 1);
    dd.insert(dd.end(), emf->d3, emf->d3 + n3 + 1);
    check(hipMemcpyAsync(dw, dd.data(), sizeof(float) * dd.size(), hipMemcpyHostToDevice, s), "h2d");
    emmi::launch_zero_c128(b_dev, n, s);
    emmi::launch_zero_c128(x_dev, n, s);
    emmi::launch_scatter_set(b_dev, di, dv, (int)nb, s);
    emmi::launch_scatter_set(x_dev, di + nb, dv + nb, (int)nx, s);
    emmi::launch_add_curl_m(b_dev, x_dev, n1, n2, n3, dw, dw + n1 + 1, dw + n1 + 1 + n2 + 1, s);
    check(hipStreamSynchronize(s), "sync");
    check(hipGetLastError(), "adjoint source");
    return 0;
  } catch (const std::exception& e) {
    g_path_err = e.what();
    return -1;
  }
}
}  // namespace

// src/inject_extract.c:124-314
extern "C" int emmi_inject_adj_source_dev(acq_t* acq, emf_t* emf, int ifreq, void* x_dev, void* b_dev, void* stream) {
  return adj_source(acq, emf, ifreq, x_dev, b_dev, stream, false);
}

// the same with the device tables in the static scratch: no allocation
extern "C" int emmi_inject_adj_source_dev_noalloc(acq_t* acq, emf_t* emf, int ifreq, void* x_dev, void* b_dev,
                                                  void* stream) {
  return adj_source(acq, emf, ifreq, x_dev, b_dev, stream, true);
}

namespace {
template <typename V>
void put(char* dev, size_t& off, const std::vector<V>& h, const V*& out, hipStream_t s) {
  off = (off + 255) & ~(size_t)255;
  check(hipMemcpyAsync(dev + off, h.data(), sizeof(V) * h.size(), hipMemcpyHostToDevice, s), "h2d");
  out = (const V*)(dev + off);
  off += sizeof(V) * h.size();
}
size_t padded(size_t b) { return (b + 255) & ~(size_t)255; }
void in_range(const std::vector<int>& v, int lo, int hi, const char* what) {
  for (int x : v)
    if (x < lo || x > hi) throw std::runtime_error(std::string("model cell centre outside the multigrid (") + what + ")");
}
}  // namespace

// src/inject_extract.c:477-635
extern "C" int emmi_extract_electric_field_dev(acq_t* acq, emf_t* emf, int ifreq, const void* x_dev, void* e_dev,
                                               void* stream) {
  (void)acq;
  (void)ifreq;
  try {
    hipStream_t s = (hipStream_t)stream;
    const emmi::CentreTables T(emf);
    const int n1 = emf->n1, n2 = emf->n2, n3 = emf->n3;
    // the gathers read corners (v, v+1) of each table entry; the vertical
    // current also the cells around the Ez edges (a-1..a+1, b-1..b+1)
    in_range(T.i1s, 0, n1 - 1, "x"); in_range(T.i2s, 0, n2 - 1, "y"); in_range(T.i3, 0, n3 - 1, "z");
    in_range(T.i1, 1, n1 - 2, "x"); in_range(T.i2, 1, n2 - 2, "y"); in_range(T.i3s, 0, n3 - 2, "z");
    const size_t ncell = (size_t)n1 * n2 * n3, nxyz = (size_t)T.nx * T.ny * T.nz;
    const size_t tb = 6 * padded(sizeof(int) * (size_t)std::max(T.nx, std::max(T.ny, T.nz))) * 2;
    char* dt = (char*)g_tab.get(tb);
    size_t off = 0;
    emmi::CentreDev D;
    put(dt, off, T.i1, D.i1, s); put(dt, off, T.i1s, D.i1s, s); put(dt, off, T.w1, D.w1, s); put(dt, off, T.w1s, D.w1s, s);
    put(dt, off, T.i2, D.i2, s); put(dt, off, T.i2s, D.i2s, s); put(dt, off, T.w2, D.w2, s); put(dt, off, T.w2s, D.w2s, s);
    put(dt, off, T.i3, D.i3, s); put(dt, off, T.i3s, D.i3s, s); put(dt, off, T.w3, D.w3, s); put(dt, off, T.w3s, D.w3s, s);
    D.nx = T.nx; D.ny = T.ny; D.nz = T.nz;
    char* dm = (char*)g_med.get(padded(sizeof(float) * ncell) + sizeof(float) * nxyz);
    check(hipMemcpyAsync(dm, &emf->sigma33[0][0][0], sizeof(float) * ncell, hipMemcpyHostToDevice, s), "h2d");
    check(hipMemcpyAsync(dm + padded(sizeof(float) * ncell), &emf->rho33[0][0][0], sizeof(float) * nxyz,
                         hipMemcpyHostToDevice, s), "h2d");
    emmi::launch_efield(D, x_dev, n1, n2, n3, (const float*)dm, (const float*)(dm + padded(sizeof(float) * ncell)),
                        e_dev, s);
    check(hipStreamSynchronize(s), "sync");
    check(hipGetLastError(), "efield");
    return 0;
  } catch (const std::exception& e) {
    g_path_err = e.what();
    return -1;
  }
}

// src/inject_extract.c:638-785
extern "C" int emmi_extract_magnetic_field_dev(acq_t* acq, emf_t* emf, int ifreq, const void* h_dev, void* e_dev,
                                               void* stream) {
  (void)acq;
  (void)ifreq;
  try {
    hipStream_t s = (hipStream_t)stream;
    const emmi::CentreTables T(emf);
    const int n1 = emf->n1, n2 = emf->n2, n3 = emf->n3;
    // every gather reads corners (v, v+1) of its table entries
    in_range(T.i1, 0, n1 - 1, "x"); in_range(T.i2, 0, n2 - 1, "y"); in_range(T.i3, 0, n3 - 1, "z");
    in_range(T.i1s, 0, n1 - 1, "x"); in_range(T.i2s, 0, n2 - 1, "y"); in_range(T.i3s, 0, n3 - 1, "z");
    const size_t tb = 6 * padded(sizeof(int) * (size_t)std::max(T.nx, std::max(T.ny, T.nz))) * 2;
    char* dt = (char*)g_tab.get(tb);
    size_t off = 0;
    emmi::CentreDev D;
    put(dt, off, T.i1, D.i1, s); put(dt, off, T.i1s, D.i1s, s); put(dt, off, T.w1, D.w1, s); put(dt, off, T.w1s, D.w1s, s);
    put(dt, off, T.i2, D.i2, s); put(dt, off, T.i2s, D.i2s, s); put(dt, off, T.w2, D.w2, s); put(dt, off, T.w2s, D.w2s, s);
    put(dt, off, T.i3, D.i3, s); put(dt, off, T.i3s, D.i3s, s); put(dt, off, T.w3, D.w3, s); put(dt, off, T.w3s, D.w3s, s);
    D.nx = T.nx; D.ny = T.ny; D.nz = T.nz;
    emmi::launch_hfield(D, h_dev, n1, n2, n3, e_dev, s);
    check(hipStreamSynchronize(s), "sync");
    check(hipGetLastError(), "hfield");
    return 0;
  } catch (const std::exception& e) {
    g_path_err = e.what();
    return -1;
  }
}

// src/build_gradient.c:13-111
extern "C" int emmi_build_gradient_dev(acq_t* acq, emf_t* emf, fwi_t* fwi, const void* const* ef_dev,
                                       const void* const* ea_dev, float* grad_dev, float* phess_dev, void* stream) {
  try {
    hipStream_t s = (hipStream_t)stream;
    if (emf->nfreq < 1 || emf->nfreq > emmi::GRAD_MAX_FREQ) throw std::runtime_error("nfreq outside 1..32");
    if (fwi->npar < 1 || fwi->npar > emmi::GRAD_MAX_PAR) throw std::runtime_error("npar outside 1..8");
    if (fwi->preco == 1 && !phess_dev) throw std::runtime_error("preco == 1 needs the pseudo-Hessian output");
    const size_t nxyz = (size_t)emf->nx * emf->ny * emf->nz, nxy = (size_t)emf->nx * emf->ny;
    emmi::GradArgs a;
    for (int f = 0; f < emf->nfreq; f++) {
      if (!ef_dev[f] || !ea_dev[f]) throw std::runtime_error("missing field of a frequency");
      a.ef[f] = ef_dev[f];
      a.ea[f] = ea_dev[f];
    }
    a.nfreq = emf->nfreq;
    a.npar = fwi->npar;
    a.preco = fwi->preco;
    for (int k = 0; k < fwi->npar; k++) a.idxpar[k] = fwi->idxpar[k];
    a.nx = emf->nx; a.ny = emf->ny; a.nz = emf->nz;
    a.s = emf->dx * emf->dy * emf->dz;
    a.dzf = emf->dz;
    a.oz = (double)emf->oz;
    a.dz = (double)emf->dz;
    char* dm = (char*)g_med.get(2 * padded(sizeof(float) * nxyz) + sizeof(float) * nxy);
    check(hipMemcpyAsync(dm, &emf->rho11[0][0][0], sizeof(float) * nxyz, hipMemcpyHostToDevice, s), "h2d");
    check(hipMemcpyAsync(dm + padded(sizeof(float) * nxyz), &emf->rho33[0][0][0], sizeof(float) * nxyz,
                         hipMemcpyHostToDevice, s), "h2d");
    check(hipMemcpyAsync(dm + 2 * padded(sizeof(float) * nxyz), &emf->bathy[0][0], sizeof(float) * nxy,
                         hipMemcpyHostToDevice, s), "h2d");
    a.rho11 = (const float*)dm;
    a.rho33 = (const float*)(dm + padded(sizeof(float) * nxyz));
    a.bathy = (const float*)(dm + 2 * padded(sizeof(float) * nxyz));
    a.grad = grad_dev;
    a.phess = phess_dev;
    emmi::launch_gradient(a, s);
    std::vector<int> boxes;
    long maxvol = 0;
    auto add = [&](float px, float py, float pz, float r) {
      int b[6];
      emmi::mute_bounds(emf, px, py, pz, r, b);
      boxes.insert(boxes.end(), b, b + 6);
      maxvol = std::max<long>(maxvol, (long)(b[1] - b[0] + 1) * (b[3] - b[2] + 1) * (b[5] - b[4] + 1));
    };
    for (int is = 0; is < acq->nsrc; is++) add(acq->src_x1[is], acq->src_x2[is], acq->src_x3[is], emf->dsmute);
    for (int ir = 0; ir < acq->nrec; ir++) add(acq->rec_x1[ir], acq->rec_x2[ir], acq->rec_x3[ir], emf->drmute);
    const int nbox = (int)(boxes.size() / 6);
    if (nbox > 65535) throw std::runtime_error("too many mute boxes");
    if (nbox > 0) {
      int* db = (int*)g_box.get(sizeof(int) * boxes.size());
      check(hipMemcpyAsync(db, boxes.data(), sizeof(int) * boxes.size(), hipMemcpyHostToDevice, s), "h2d");
      emmi::launch_mute(grad_dev, fwi->npar, emf->nx, emf->ny, emf->nz, db, nbox, (int)maxvol, s);
    }
    check(hipStreamSynchronize(s), "sync");
    check(hipGetLastError(), "gradient");
    return 0;
  } catch (const std::exception& e) {
    g_path_err = e.what();
    return -1;
  }
}

extern "C" const char* emmi_path_last_error(void) { return g_path_err.c_str(); }
