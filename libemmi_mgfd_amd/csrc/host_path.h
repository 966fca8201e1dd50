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
