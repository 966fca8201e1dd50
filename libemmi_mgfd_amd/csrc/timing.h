// timing.h -- optional per-kernel-class timing with HIP events recorded on
// the stream each kernel is launched on.
#pragma once

#include <hip/hip_runtime.h>

namespace emmi {
namespace timing {

enum Class { RESIDUAL = 0, LINE = 1, RESTRICT = 2, PROLONG = 3, OTHER = 4, LINE_RHS = 5, NCLASS = 6 };

void enable(bool on);
bool enabled();
// bracket one launch; no-ops when disabled
void begin(int cls, hipStream_t s);
void end(int cls, hipStream_t s, double bytes);
// events for a launch that records them in its own dispatch
// (hipExtLaunchKernelGGL), so the elapsed time is the kernel's alone; both
// null when disabled
void kernel_events(int cls, double bytes, hipEvent_t* start, hipEvent_t* stop);
// synchronises the recorded events, returns accumulated ms, launch counts and
// algorithmic bytes per class
int read(double* ms, long* count, double* bytes, int n);

struct Scope {
  int cls;
  hipStream_t s;
  double bytes;
  Scope(int c, hipStream_t st, double b = 0) : cls(c), s(st), bytes(b) { begin(cls, s); }
  ~Scope() { end(cls, s, bytes); }
};

}  // namespace timing
}  // namespace emmi
