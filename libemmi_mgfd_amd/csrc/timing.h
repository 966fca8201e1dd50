// timing.h -- optional per-kernel-class timing with HIP events recorded on
// the stream each kernel is launched on.
#pragma once

#include <hip/hip_runtime.h>

namespace emmi {
namespace timing {

enum Class { RESIDUAL = 0, LINE = 1, RESTRICT = 2, PROLONG = 3, OTHER = 4, NCLASS = 5 };

void enable(bool on);
bool enabled();
// bracket one launch; no-ops when disabled
void begin(int cls, hipStream_t s);
void end(int cls, hipStream_t s);
// synchronises the recorded events, returns accumulated ms and counts
int read(double* ms, long* count, int n);

struct Scope {
  int cls;
  hipStream_t s;
  Scope(int c, hipStream_t st) : cls(c), s(st) { begin(cls, s); }
  ~Scope() { end(cls, s); }
};

}  // namespace timing
}  // namespace emmi
