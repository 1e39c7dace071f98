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

// profiler events still pending (called before a stream is destroyed)
void prof_fold_pending() {
  std::lock_guard<std::mutex> g(mu);
  fold();
}

ProfScope::ProfScope(hipStream_t s, const char* name, double bytes, double flops)
    : s_(s), name_(name), bytes_(bytes), flops_(flops) {
  if (!enabled || !name) return;  // a null name records nothing
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

// One staging ring of pinned host memory per host thread.  Every host-to-
// device copy staged through it records an event on its stream right after the
// copy is queued (one event per stream, re-recorded); a ring that wraps waits
// for those events only, never for the whole device, so threads that drive
// other streams are not held up.  Device-to-host copies land in the ring and
// wait for their own stream.  A ring whose buffer is too small for one copy
// moves to a larger buffer and keeps the old one (copies may still read it)
// until the process ends.  Threads that end hand their ring back with its
// pending events; the next owner waits for them before it writes.
namespace {
struct Mark {
  hipStream_t s;
  hipEvent_t e;
};
struct Ring {
  char* base = nullptr;
  size_t cap = 0, off = 0;
  std::vector<Mark> marks;  // last copy per stream since the last wrap
  std::vector<hipEvent_t> spare_ev;
  std::vector<char*> retired;
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

Ring& ring() {
  if (!my_ring.r) {
    std::lock_guard<std::mutex> g(ring_mu);
    if (!spare_rings.empty()) {
      my_ring.r = spare_rings.back();
      spare_rings.pop_back();
      my_ring.r->off = my_ring.r->cap;  // wait for the old owner's copies first
    } else {
      my_ring.r = new Ring();
    }
  }
  return *my_ring.r;
}

void wait_marks(Ring& R) {
  for (auto& m : R.marks) {
    HIP_OK(hipEventSynchronize(m.e));
    R.spare_ev.push_back(m.e);
  }
  R.marks.clear();
}

// a region of `bytes` of pinned memory that no queued copy still reads
char* take(Ring& R, size_t bytes) {
  const size_t need = (bytes + 255) & ~size_t(255);
  if (!R.base || need > R.cap) {
    if (R.base) R.retired.push_back(R.base);
    for (auto& m : R.marks) R.spare_ev.push_back(m.e);  // their copies read the old buffer
    R.marks.clear();
    R.cap = std::max<size_t>(size_t(64) << 20, need);
    HIP_OK(hipHostMalloc((void**)&R.base, R.cap, hipHostMallocDefault));
    R.off = 0;
  }
  if (R.off + need > R.cap) {
    wait_marks(R);
    R.off = 0;
  }
  char* p = R.base + R.off;
  R.off += need;
  return p;
}

void mark(Ring& R, hipStream_t s) {
  for (auto& m : R.marks)
    if (m.s == s) {
      HIP_OK(hipEventRecord(m.e, s));
      return;
    }
  hipEvent_t e;
  if (!R.spare_ev.empty()) {
    e = R.spare_ev.back();
    R.spare_ev.pop_back();
  } else {
    HIP_OK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
  }
  HIP_OK(hipEventRecord(e, s));
  R.marks.push_back(Mark{s, e});
}

constexpr size_t kChunk = size_t(16) << 20;
}  // namespace

// Streams go back to a pool instead of being destroyed, so that no event a
// ring or the profiler still holds refers to a destroyed stream.
namespace {
std::mutex pool_mu;
std::map<unsigned, std::vector<hipStream_t>> stream_pool;
std::map<hipStream_t, unsigned> stream_flags;
}  // namespace

hipStream_t stream_create(unsigned flags) {
  {
    std::lock_guard<std::mutex> g(pool_mu);
    auto& v = stream_pool[flags];
    if (!v.empty()) {
      hipStream_t s = v.back();
      v.pop_back();
      return s;
    }
  }
  hipStream_t s = nullptr;
  HIP_OK(hipStreamCreateWithFlags(&s, flags));
  std::lock_guard<std::mutex> g(pool_mu);
  stream_flags[s] = flags;
  return s;
}

void stream_destroy(hipStream_t s) {
  if (!s) return;
  (void)hipStreamSynchronize(s);
  prof_fold_pending();
  std::lock_guard<std::mutex> g(pool_mu);
  stream_pool[stream_flags[s]].push_back(s);
}

void h2d(void* dst, const void* src, size_t bytes, hipStream_t s) {
  Ring& R = ring();
  for (size_t o = 0; o < bytes; o += kChunk) {
    const size_t n = std::min(kChunk, bytes - o);
    char* p = take(R, n);
    std::memcpy(p, (const char*)src + o, n);
    HIP_OK(hipMemcpyAsync((char*)dst + o, p, n, hipMemcpyHostToDevice, s));
    mark(R, s);
  }
}

void h2d2(void* dst, const void* a, size_t na, const void* b, size_t nb, hipStream_t s) {
  const size_t oa = (na + 255) & ~size_t(255);
  if (oa + nb > kChunk) {
    h2d(dst, a, na, s);
    h2d((char*)dst + oa, b, nb, s);
    return;
  }
  Ring& R = ring();
  char* p = take(R, oa + nb);
  std::memcpy(p, a, na);
  std::memcpy(p + oa, b, nb);
  HIP_OK(hipMemcpyAsync(dst, p, oa + nb, hipMemcpyHostToDevice, s));
  mark(R, s);
}

void d2h(void* dst, const void* src, size_t bytes, hipStream_t s) {
  Ring& R = ring();
  for (size_t o = 0; o < bytes; o += kChunk) {
    const size_t n = std::min(kChunk, bytes - o);
    char* p = take(R, n);
    HIP_OK(hipMemcpyAsync(p, (const char*)src + o, n, hipMemcpyDeviceToHost, s));
    HIP_OK(hipStreamSynchronize(s));
    std::memcpy((char*)dst + o, p, n);
  }
  if (!bytes) HIP_OK(hipStreamSynchronize(s));
}

}  // namespace aifmm
