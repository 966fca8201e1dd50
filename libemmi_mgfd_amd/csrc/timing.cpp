// timing.cpp -- event pairs per launch, resolved when the counters are read.
#include "timing.h"

#include <mutex>
#include <vector>

namespace emmi {
namespace timing {
namespace {
struct Pending {
  int cls;
  hipEvent_t a, b;
  double bytes, layout;
};
bool g_on = false;
std::mutex g_mu;
std::vector<Pending> g_pending;
std::vector<hipEvent_t> g_pool;
hipEvent_t g_open[NCLASS];
double g_ms[NCLASS];
long g_count[NCLASS];
double g_bytes[NCLASS];
double g_layout[NCLASS];

hipEvent_t get_event() {
  if (!g_pool.empty()) {
    hipEvent_t e = g_pool.back();
    g_pool.pop_back();
    return e;
  }
  hipEvent_t e;
  hipEventCreate(&e);
  return e;
}
}  // namespace

void enable(bool on) {
  std::lock_guard<std::mutex> lk(g_mu);
  g_on = on;
  for (int i = 0; i < NCLASS; i++) { g_ms[i] = 0; g_count[i] = 0; g_bytes[i] = 0; g_layout[i] = 0; }
}

bool enabled() { return g_on; }

void begin(int cls, hipStream_t s) {
  if (!g_on) return;
  std::lock_guard<std::mutex> lk(g_mu);
  g_open[cls] = get_event();
  hipEventRecord(g_open[cls], s);
}

void end(int cls, hipStream_t s, double bytes, double layout_bytes) {
  if (!g_on) return;
  std::lock_guard<std::mutex> lk(g_mu);
  hipEvent_t e = get_event();
  hipEventRecord(e, s);
  g_pending.push_back({cls, g_open[cls], e, bytes, layout_bytes < 0 ? bytes : layout_bytes});
}

void kernel_events(int cls, double bytes, hipEvent_t* start, hipEvent_t* stop, double layout_bytes) {
  *start = nullptr;
  *stop = nullptr;
  if (!g_on) return;
  std::lock_guard<std::mutex> lk(g_mu);
  *start = get_event();
  *stop = get_event();
  g_pending.push_back({cls, *start, *stop, bytes, layout_bytes < 0 ? bytes : layout_bytes});
}

int read(double* ms, long* count, double* bytes, int n) {
  std::lock_guard<std::mutex> lk(g_mu);
  for (auto& p : g_pending) {
    hipEventSynchronize(p.b);
    float t = 0;
    hipEventElapsedTime(&t, p.a, p.b);
    g_ms[p.cls] += t;
    g_count[p.cls] += 1;
    g_bytes[p.cls] += p.bytes;
    g_layout[p.cls] += p.layout;
    g_pool.push_back(p.a);
    g_pool.push_back(p.b);
  }
  g_pending.clear();
  for (int i = 0; i < n && i < NCLASS; i++) {
    ms[i] = g_ms[i];
    count[i] = g_count[i];
    if (bytes) bytes[i] = g_bytes[i];
  }
  return NCLASS;
}

int read_layout(double* bytes, int n) {
  std::lock_guard<std::mutex> lk(g_mu);
  for (int i = 0; i < n && i < NCLASS; i++) bytes[i] = g_layout[i];
  return NCLASS;
}

}  // namespace timing
}  // namespace emmi
