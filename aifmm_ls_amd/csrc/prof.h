// Per-kernel timing with HIP events recorded on the launching stream, plus
// the algorithmic bytes and flops each launch was given by its host caller.
// Off by default; when on, every launch made through a ProfScope records an
// event pair, and prof_collect() sums elapsed times per kernel name.
#pragma once
#include <string>
#include <vector>

#include "common.h"

namespace aifmm {

struct KernelProfile {
  std::string name;
  long long launches = 0;
  double ms = 0;      // summed event-to-event duration
  double bytes = 0;   // algorithmic bytes (operands read + results written)
  double flops = 0;   // algorithmic flops (complex FMA = 8)
};

// log_launches: also keep one record per launch (name, bytes, duration)
void prof_enable(bool on, bool log_launches = false);
bool prof_on();
// Synchronises on the recorded events and folds them into the totals.
std::vector<KernelProfile> prof_collect(bool reset);

struct LaunchRecord {
  const char* name;
  double bytes, ms;
};
// The per-launch records in launch order (after folding the pending events).
std::vector<LaunchRecord> prof_launches(bool reset);

class ProfScope {
 public:
  ProfScope(hipStream_t s, const char* name, double bytes, double flops);
  ~ProfScope();

 private:
  hipStream_t s_;
  const char* name_;
  double bytes_, flops_;
  hipEvent_t e0_ = nullptr;
};

}  // namespace aifmm
