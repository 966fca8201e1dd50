// timing.h -- optional per-kernel-class timing with HIP events recorded on
// the stream each kernel is launched on.
#pragma once

#include <hip/hip_runtime.h>

namespace emmi {
namespace timing {

// CONV / CONV_DW: the initial-guess network's convolutions (forward and
// input gradient / weight gradient); their "bytes" are algorithmic flops,
// 2 * 27 * Cin * Cout per voxel
enum Class { RESIDUAL = 0, LINE = 1, RESTRICT = 2, PROLONG = 3, OTHER = 4, LINE_RHS = 5, CONV = 6, CONV_DW = 7,
             NCLASS = 8 };

// Every launch carries two byte counts: its compulsory bytes (each input read
// once and each output written once at its level's shape; a full line sweep,
// all four colours, is 3 Ne s + 3 Nc c, one colour a quarter of that) and the
// bytes of its own data layout (for the line kernels: their stored factor and
// line records as well; for the other kernels the same as the first).
void enable(bool on);
bool enabled();
// bracket one launch; no-ops when disabled
void begin(int cls, hipStream_t s);
void end(int cls, hipStream_t s, double bytes, double layout_bytes = -1.0);
// events for a launch that records them in its own dispatch
// (hipExtLaunchKernelGGL), so the elapsed time is the kernel's alone; both
// null when disabled
void kernel_events(int cls, double bytes, hipEvent_t* start, hipEvent_t* stop, double layout_bytes = -1.0);
// synchronises the recorded events, returns accumulated ms, launch counts and
// compulsory bytes per class
int read(double* ms, long* count, double* bytes, int n);
// the layout bytes per class accumulated by the last read
int read_layout(double* bytes, int n);

struct Scope {
  int cls;
  hipStream_t s;
  double bytes, layout;
  Scope(int c, hipStream_t st, double b = 0, double lb = -1.0) : cls(c), s(st), bytes(b), layout(lb) { begin(cls, s); }
  ~Scope() { end(cls, s, bytes, layout); }
};

}  // namespace timing
}  // namespace emmi
