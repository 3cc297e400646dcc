// Per-launch HIP-event timing of one kernel class (see common.h): events
// recorded around the launch, or stamped on the dispatch by hipExtLaunchKernel.
#include "common.h"

#include <vector>

namespace msdp {

namespace {
struct Profiler {
  int cls = PROF_NONE;
  hipStream_t stream = nullptr;
  std::vector<hipEvent_t> pool;
  size_t used = 0;
  std::vector<double> work;
};
Profiler g_prof;

hipEvent_t next_event() {
  if (g_prof.used == g_prof.pool.size()) {
    hipEvent_t e;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    g_prof.pool.push_back(e);
  }
  return g_prof.pool[g_prof.used++];
}
}  // namespace

bool prof_active(int cls, hipStream_t s) { return g_prof.cls == cls && g_prof.stream == s; }

void prof_before(hipStream_t s) {
  hipEvent_t e = next_event();
  if (e) (void)hipEventRecord(e, s);
}

void prof_after(hipStream_t s, double work) {
  hipEvent_t e = next_event();
  if (e) (void)hipEventRecord(e, s);
  g_prof.work.push_back(work);
}

void prof_dispatch_events(hipEvent_t* start, hipEvent_t* stop, double work) {
  *start = next_event();
  *stop = next_event();
  g_prof.work.push_back(work);
}

int prof_begin(int cls, hipStream_t s) {
  g_prof.cls = cls;
  g_prof.stream = s;
  g_prof.used = 0;
  g_prof.work.clear();
  return 0;
}

int prof_end(double* total_ms, int64_t* launches, double* work) {
  if (g_prof.stream) MSDP_HIP(hipStreamSynchronize(g_prof.stream));
  double ms = 0.0, w = 0.0;
  const size_t n = g_prof.work.size();
  for (size_t k = 0; k < n && 2 * k + 1 < g_prof.used; ++k) {
    float t = 0.f;
    MSDP_HIP(hipEventElapsedTime(&t, g_prof.pool[2 * k], g_prof.pool[2 * k + 1]));
    ms += t;
    w += g_prof.work[k];
  }
  *total_ms = ms;
  *launches = (int64_t)n;
  *work = w;
  g_prof.cls = PROF_NONE;
  g_prof.stream = nullptr;
  return 0;
}

}  // namespace msdp
