// Per-launch HIP-event timing of one kernel class (see common.h): events
// recorded around the launch, or stamped on the dispatch by hipExtLaunchKernel.
// Either for one stream (a single solve) or for every stream of the process
// (the serving loop, whose worker threads launch concurrently: the event pool
// and the work list are guarded by a mutex then).
#include "common.h"

#include <mutex>
#include <vector>

namespace msdp {

namespace {
struct Profiler {
  int cls = PROF_NONE;
  bool all_streams = false;
  hipStream_t stream = nullptr;
  std::vector<hipEvent_t> pool;
  size_t used = 0;
  // per launch: index of its start event (the stop event follows), work
  std::vector<size_t> start;
  std::vector<double> work;
  std::mutex mu;
};
Profiler g_prof;

hipEvent_t next_event_locked() {
  if (g_prof.used == g_prof.pool.size()) {
    hipEvent_t e;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    g_prof.pool.push_back(e);
  }
  return g_prof.pool[g_prof.used++];
}

thread_local size_t t_pending = 0;  // start event of this thread's open launch
}  // namespace

bool prof_active(int cls, hipStream_t s) {
  return g_prof.cls == cls && (g_prof.all_streams || g_prof.stream == s);
}

void prof_before(hipStream_t s) {
  hipEvent_t e;
  {
    std::lock_guard<std::mutex> lk(g_prof.mu);
    t_pending = g_prof.used;
    e = next_event_locked();
    (void)next_event_locked();  // its stop event
  }
  if (e) (void)hipEventRecord(e, s);
}

void prof_after(hipStream_t s, double work) {
  hipEvent_t e;
  {
    std::lock_guard<std::mutex> lk(g_prof.mu);
    if (t_pending + 1 >= g_prof.pool.size()) return;
    e = g_prof.pool[t_pending + 1];
    g_prof.start.push_back(t_pending);
    g_prof.work.push_back(work);
  }
  (void)hipEventRecord(e, s);
}

void prof_dispatch_events(hipEvent_t* start, hipEvent_t* stop, double work) {
  std::lock_guard<std::mutex> lk(g_prof.mu);
  const size_t k = g_prof.used;
  *start = next_event_locked();
  *stop = next_event_locked();
  g_prof.start.push_back(k);
  g_prof.work.push_back(work);
}

int prof_begin(int cls, hipStream_t s) {
  std::lock_guard<std::mutex> lk(g_prof.mu);
  g_prof.cls = cls;
  g_prof.all_streams = s == nullptr;
  g_prof.stream = s;
  g_prof.used = 0;
  g_prof.start.clear();
  g_prof.work.clear();
  return 0;
}

int prof_end(double* total_ms, int64_t* launches, double* work) {
  if (g_prof.all_streams) MSDP_HIP(hipDeviceSynchronize());
  else if (g_prof.stream) MSDP_HIP(hipStreamSynchronize(g_prof.stream));
  std::lock_guard<std::mutex> lk(g_prof.mu);
  double ms = 0.0, w = 0.0;
  const size_t n = g_prof.work.size();
  for (size_t k = 0; k < n; ++k) {
    const size_t i = g_prof.start[k];
    if (i + 1 >= g_prof.used) continue;
    float t = 0.f;
    MSDP_HIP(hipEventElapsedTime(&t, g_prof.pool[i], g_prof.pool[i + 1]));
    ms += t;
    w += g_prof.work[k];
  }
  *total_ms = ms;
  *launches = (int64_t)n;
  *work = w;
  g_prof.cls = PROF_NONE;
  g_prof.all_streams = false;
  g_prof.stream = nullptr;
  return 0;
}

}  // namespace msdp
