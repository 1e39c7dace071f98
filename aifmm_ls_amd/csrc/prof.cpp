#include "prof.h"
#include "batch.h"

#include <algorithm>
#include <cstring>
#include <map>
#include <mutex>

namespace aifmm {

namespace {
struct Pending {
  const char* name;
  double bytes, flops;
  hipEvent_t e0, e1;
};
std::mutex mu;
bool enabled = false, log_on = false;
std::vector<LaunchRecord> launch_log;
std::vector<Pending> pending;
std::vector<hipEvent_t> pool;
std::map<std::string, KernelProfile> totals;

hipEvent_t take() {
  if (!pool.empty()) {
    hipEvent_t e = pool.back();
    pool.pop_back();
    return e;
  }
  hipEvent_t e;
  HIP_OK(hipEventCreate(&e));
  return e;
}

void fold() {
  for (auto& p : pending) {
    HIP_OK(hipEventSynchronize(p.e1));
    float ms = 0;
    HIP_OK(hipEventElapsedTime(&ms, p.e0, p.e1));
    KernelProfile& k = totals[p.name];
    k.name = p.name;
    k.launches += 1;
    k.ms += ms;
    k.bytes += p.bytes;
    k.flops += p.flops;
    if (log_on) launch_log.push_back(LaunchRecord{p.name, p.bytes, (double)ms});
    pool.push_back(p.e0);
    pool.push_back(p.e1);
  }
  pending.clear();
}
}  // namespace

void prof_enable(bool on, bool log_launches) {
  std::lock_guard<std::mutex> g(mu);
  enabled = on;
  log_on = on && log_launches;
}

bool prof_on() { return enabled; }

std::vector<KernelProfile> prof_collect(bool reset) {
  std::lock_guard<std::mutex> g(mu);
  fold();
  std::vector<KernelProfile> out;
  for (auto& t : totals) out.push_back(t.second);
  if (reset) totals.clear();
  return out;
}

std::vector<LaunchRecord> prof_launches(bool reset) {
  std::lock_guard<std::mutex> g(mu);
  fold();
  std::vector<LaunchRecord> out = launch_log;
  if (reset) launch_log.clear();
  return out;
}

ProfScope::ProfScope(hipStream_t s, const char* name, double bytes, double flops)
    : s_(s), name_(name), bytes_(bytes), flops_(flops) {
  if (!enabled) return;
  std::lock_guard<std::mutex> g(mu);
  e0_ = take();
  HIP_OK(hipEventRecord(e0_, s_));
}

ProfScope::~ProfScope() {
  if (!e0_) return;
  std::lock_guard<std::mutex> g(mu);
  hipEvent_t e1 = take();
  if (hipEventRecord(e1, s_) != hipSuccess) return;
  pending.push_back(Pending{name_, bytes_, flops_, e0_, e1});
  // bound the number of outstanding events
  if (pending.size() > 8192) fold();
}

// One staging ring per host thread.  A thread queues the copy out of its
// region before it asks for the next one, so after a device synchronise every
// copy the thread queued from its ring has completed: a ring that wraps
// synchronises and starts over.  Threads that end hand their ring back; the
// next owner starts with a wrap, since the old owner's copies may be queued.
namespace {
struct Ring {
  char* base = nullptr;
  size_t cap = 0, off = 0;
};
std::mutex ring_mu;
std::vector<Ring*> spare_rings;
struct RingSlot {
  Ring* r = nullptr;
  ~RingSlot() {
    if (!r) return;
    std::lock_guard<std::mutex> g(ring_mu);
    spare_rings.push_back(r);
  }
};
thread_local RingSlot my_ring;
}  // namespace

void* stage_pinned(const void* src, size_t bytes, hipStream_t s) {
  (void)s;
  if (!my_ring.r) {
    std::lock_guard<std::mutex> g(ring_mu);
    if (!spare_rings.empty()) {
      my_ring.r = spare_rings.back();
      spare_rings.pop_back();
      my_ring.r->off = my_ring.r->cap;  // force a synchronised wrap first
    } else {
      my_ring.r = new Ring();
    }
  }
  Ring& R = *my_ring.r;
  const size_t need = (bytes + 255) & ~size_t(255);
  if (!R.base || need > R.cap) {
    if (R.base) {
      HIP_OK(hipDeviceSynchronize());
      HIP_OK(hipHostFree(R.base));
    }
    R.cap = std::max<size_t>(size_t(64) << 20, need);
    HIP_OK(hipHostMalloc((void**)&R.base, R.cap, hipHostMallocDefault));
    R.off = 0;
  }
  if (R.off + need > R.cap) {
    HIP_OK(hipDeviceSynchronize());
    R.off = 0;
  }
  void* p = R.base + R.off;
  std::memcpy(p, src, bytes);
  R.off += need;
  return p;
}

}  // namespace aifmm
