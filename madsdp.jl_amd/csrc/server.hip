// Serving loop: continuous batching of independent SDP solves on one device.
//
// Each of `slots` worker threads owns a HIP stream and takes the first request
// (in submission order) that has arrived and whose memory footprint fits in
// the budget next to the solves already running (a request that does not fit
// is passed over for at most HOLD_S, then waits at the head); the solve is uploaded, run
// and released on the slot's stream, so the admitted solves overlap on the
// device while requests that arrive meanwhile wait for a slot and memory.
// Every kernel of a solve is deterministic and runs on that solve's stream
// only, so a result does not depend on what else is in the mix.
#include "common.h"
#include "madsdp_hip.h"

#include <chrono>
#include <condition_variable>
#include <deque>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

namespace {

struct Req {
  msdp_solver* s = nullptr;
  double arrival = 0, admitted = -1, uploaded = -1, done = -1;
  int64_t footprint = 0;
  int rc = 0, slot = -1;
  bool finished = false;
  std::string err;
};

}  // namespace

// admission: how far past the oldest waiting request first fit may look,
// and how long a request that does not fit may be passed over
constexpr size_t ADMIT_WINDOW = 256;
constexpr double HOLD_S = 1.0;

struct msdp_server {
  int device = 0, slots = 1;
  int64_t budget = 0, in_use = 0, peak_bytes = 0;
  int active = 0, peak_active = 0;
  std::vector<hipStream_t> streams;
  std::vector<std::thread> workers;
  std::mutex mu;
  std::condition_variable cv;
  std::deque<Req> reqs;  // references stay valid while requests are appended
  size_t next = 0;
  int64_t completed = 0;
  bool started = false, closing = false;
  bool blocking = true;  // worker threads sleep while their solve's kernels run
  std::chrono::steady_clock::time_point t0;

  double elapsed() const {
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  }
  void work(int slot);
};

void msdp_server::work(int slot) {
  (void)hipSetDevice(device);
  msdp::set_blocking_sync(blocking);
  for (;;) {
    Req* r = nullptr;
    {
      std::unique_lock<std::mutex> lk(mu);
      for (;;) {
        if (next < reqs.size()) {
          // first fit in submission order among the requests that have
          // arrived: a request whose footprint does not fit yet is passed
          // over by later, smaller ones, unless it has waited HOLD_S already
          // (then it is admitted first, as soon as memory frees)
          const double now = elapsed();
          double wait_s = -1.0;
          for (size_t i = next; i < reqs.size() && i < next + ADMIT_WINDOW; ++i) {
            Req& h = reqs[i];
            if (h.admitted >= 0) continue;
            if (now < h.arrival) {
              wait_s = h.arrival - now;
              break;
            }
            if (active == 0 || in_use + h.footprint <= budget) {
              r = &h;
              break;
            }
            if (now - h.arrival > HOLD_S) break;
          }
          if (r) break;
          if (wait_s > 0) cv.wait_for(lk, std::chrono::duration<double>(wait_s));
          else cv.wait(lk);  // memory frees when a running solve finishes
          continue;
        }
        if (closing) return;
        cv.wait(lk);
      }
      in_use += r->footprint;
      ++active;
      peak_bytes = std::max(peak_bytes, in_use);
      peak_active = std::max(peak_active, active);
      r->admitted = elapsed();
      r->slot = slot;
      while (next < reqs.size() && reqs[next].admitted >= 0) ++next;
    }
    int rc = msdp_solver_set_stream(r->s, streams[slot]);
    if (rc == 0) rc = msdp_solver_upload(r->s);
    const double t_up = elapsed();
    if (rc == 0) rc = msdp_solver_solve(r->s);
    std::string err = rc < 0 ? std::string(msdp_last_error()) : std::string();
    // always: frees the device state and drops the slot's stream
    const int rc2 = msdp_solver_release(r->s);
    if (rc == 0 && rc2 != 0) {
      rc = rc2;
      err = msdp_last_error();
    }
    {
      std::lock_guard<std::mutex> lk(mu);
      r->rc = rc;
      r->err = err;
      r->uploaded = t_up;
      r->done = elapsed();
      r->finished = true;
      in_use -= r->footprint;
      --active;
      ++completed;
    }
    cv.notify_all();
  }
}

extern "C" {

int msdp_server_create(msdp_server** out, int device, int32_t slots, int64_t mem_budget_bytes) {
  if (!out) { msdp::set_error("null out"); return -2; }
  *out = nullptr;
  if (slots < 1 || slots > 64) { msdp::set_error("slots must be in 1..64"); return -2; }
  if (mem_budget_bytes <= 0) { msdp::set_error("memory budget must be positive"); return -2; }
  int ndev = 0;
  MSDP_HIP(hipGetDeviceCount(&ndev));
  if (device < 0 || device >= ndev) { msdp::set_error("no such HIP device"); return -2; }
  MSDP_HIP(hipSetDevice(device));
  auto* v = new msdp_server();
  v->device = device;
  v->slots = slots;
  v->budget = mem_budget_bytes;
  for (int k = 0; k < slots; ++k) {
    hipStream_t st;
    hipError_t e = hipStreamCreateWithFlags(&st, hipStreamNonBlocking);
    if (e != hipSuccess) {
      for (hipStream_t q : v->streams) (void)hipStreamDestroy(q);
      delete v;
      return msdp::fail_hip(e, "hipStreamCreateWithFlags", __FILE__, __LINE__);
    }
    v->streams.push_back(st);
  }
  *out = v;
  return 0;
}

int msdp_server_start(msdp_server* v) {
  if (!v) { msdp::set_error("null server"); return -2; }
  std::lock_guard<std::mutex> lk(v->mu);
  if (v->started) { msdp::set_error("server already started"); return -2; }
  v->started = true;
  v->t0 = std::chrono::steady_clock::now();
  for (int k = 0; k < v->slots; ++k) v->workers.emplace_back(&msdp_server::work, v, k);
  return 0;
}

int msdp_server_submit(msdp_server* v, msdp_solver* s, double arrival_s, int64_t* id) {
  if (!v || !s) { msdp::set_error("null server or solver"); return -2; }
  int64_t fp = 0;
  MSDP_TRY(msdp_solver_footprint(s, &fp));
  {
    std::lock_guard<std::mutex> lk(v->mu);
    if (v->closing) { msdp::set_error("server is shutting down"); return -2; }
    Req r;
    r.s = s;
    r.arrival = arrival_s;
    r.footprint = fp;
    v->reqs.push_back(r);
    if (id) *id = (int64_t)v->reqs.size() - 1;
  }
  v->cv.notify_all();
  return 0;
}

int msdp_server_wait_all(msdp_server* v) {
  if (!v) { msdp::set_error("null server"); return -2; }
  std::unique_lock<std::mutex> lk(v->mu);
  if (!v->started) { msdp::set_error("server not started"); return -2; }
  v->cv.wait(lk, [&] { return v->completed == (int64_t)v->reqs.size(); });
  for (const Req& r : v->reqs)
    if (r.rc < 0) { msdp::set_error("a request failed: " + r.err); return -2; }
  return 0;
}

int msdp_server_request(const msdp_server* cv_, int64_t id, msdp_request_info* out) {
  auto* v = const_cast<msdp_server*>(cv_);
  if (!v || !out) { msdp::set_error("null server or out"); return -2; }
  std::lock_guard<std::mutex> lk(v->mu);
  if (id < 0 || id >= (int64_t)v->reqs.size()) { msdp::set_error("no such request"); return -2; }
  const Req& r = v->reqs[id];
  out->arrival = r.arrival;
  out->admitted = r.admitted;
  out->uploaded = r.uploaded;
  out->done = r.done;
  out->footprint = r.footprint;
  out->rc = r.rc;
  out->slot = r.slot;
  out->finished = r.finished ? 1 : 0;
  out->pad = 0;
  return 0;
}

int msdp_server_get_stats(const msdp_server* cv_, msdp_server_stats* out) {
  auto* v = const_cast<msdp_server*>(cv_);
  if (!v || !out) { msdp::set_error("null server or out"); return -2; }
  std::lock_guard<std::mutex> lk(v->mu);
  out->submitted = (int64_t)v->reqs.size();
  out->completed = v->completed;
  out->peak_bytes = v->peak_bytes;
  out->slots = v->slots;
  out->peak_active = v->peak_active;
  out->elapsed = v->started ? v->elapsed() : 0.0;
  return 0;
}

int msdp_server_destroy(msdp_server* v) {
  if (!v) return 0;
  {
    std::lock_guard<std::mutex> lk(v->mu);
    v->closing = true;
    v->next = v->reqs.size();  // requests not yet admitted are dropped
  }
  v->cv.notify_all();
  for (auto& t : v->workers) t.join();
  for (hipStream_t q : v->streams) (void)hipStreamDestroy(q);
  delete v;
  return 0;
}

}  // extern "C"
