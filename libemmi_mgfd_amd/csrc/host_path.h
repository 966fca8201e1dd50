// host_path.h -- host-side helpers shared by the reference entry points and
// the device-resident API (cell location, channel codes, interpolation tables).
#pragma once

#include <vector>

#include "emmi_mgfd.h"

namespace emmi {

// cell (i,j,k) and trilinear weights of a point on one staggering pattern
// (0 Ex, 1 Ey, 2 Ez, 3 Hx, 4 Hy, 5 Hz) plus the float volume of that cell
struct Tri {
  int i, j, k;
  float w1, w2, w3, vol;
};
Tri locate(const emf_t* emf, int pattern, float px, float py, float pz);
int channel_code(const char* ch);   // Ex..Hz -> 0..5, -1 otherwise
int find_index_guarded(int n, const float* x, int len, float val);

struct CentreTables {
  int nx, ny, nz;
  std::vector<int> i1, i1s, i2, i2s, i3, i3s;
  std::vector<float> w1, w1s, w2, w2s, w3, w3s;
  explicit CentreTables(const emf_t* emf);
};

}  // namespace emmi

namespace emmi {
// the nonzero entries of the source term of src/inject_extract.c:26-120,
// accumulated in the reference's order (bit-identical to inject_source)
void source_entries(acq_t* acq, emf_t* emf, int ifreq, std::vector<long>& idx, std::vector<emmi_c128>& val);
// the E (ix) and H (ib) vector entries receiver extraction reads, in reading order
void receiver_indices(acq_t* acq, emf_t* emf, std::vector<long>& ix, std::vector<long>& ib);
// extract_emf_data from previously gathered entries (same order as receiver_indices)
void receiver_data_from(acq_t* acq, emf_t* emf, int ifreq, const std::vector<emmi_c128>& vx,
                        const std::vector<emmi_c128>& vb);
// the nonzero entries of the adjoint source's scatters (src/inject_extract.c:
// 133-302): b (E channels) and the magnetic dipole M (H channels), accumulated
// in the reference's order; b += curl M follows on the device
void adj_entries(acq_t* acq, emf_t* emf, int ifreq, std::vector<long>& bidx, std::vector<emmi_c128>& bval,
                 std::vector<long>& xidx, std::vector<emmi_c128>& xval);
// gradient mute box around a point: b = i1min, i1max, i2min, i2max, i3min, i3max
void mute_bounds(const emf_t* emf, float px, float py, float pz, float r, int* b);
}  // namespace emmi
