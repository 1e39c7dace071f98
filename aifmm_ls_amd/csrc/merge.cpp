// Plain C++ (no HIP): the lockstep logic is tested on the host.
#include "merge.h"

#include <chrono>

namespace aifmm {

static double now_s() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

namespace {
thread_local MergeGroup* tl_group = nullptr;
}

MergeGroup* merge_group() { return tl_group; }

MergeMember::MergeMember(MergeGroup* g) : g_(g) { tl_group = g; }

MergeMember::~MergeMember() {
  tl_group = nullptr;
  if (g_) g_->leave();
}

void MergeGroup::submit(MergeFn fn, void* batch) {
  std::unique_lock<std::mutex> lk(mu_);
  const unsigned long mine = round_;
  ops_.emplace_back(fn, batch);
  if (++arrived_ == active_) execute(lk);
  else cv_.wait(lk, [&] { return round_ != mine; });
  // the round's failures stay listed until its last member has looked
  for (auto& f : failed_)
    if (f.first == batch) {
      std::exception_ptr e = f.second;
      std::rethrow_exception(e);
    }
}

void MergeGroup::leave() {
  std::unique_lock<std::mutex> lk(mu_);
  --active_;
  if (arrived_ > 0 && arrived_ == active_) execute(lk);
}

// Called by the member that completes a round, with the lock held.  The
// submitted batches belong to members blocked in submit(), so they stay alive
// until the round ends.
void MergeGroup::execute(std::unique_lock<std::mutex>& lk) {
  std::vector<std::pair<MergeFn, void*>> ops;
  ops.swap(ops_);
  const double t0 = now_s();
  if (t_end_ >= 0) gather_s_ += t0 - t_end_;
  lk.unlock();
  std::vector<std::pair<void*, std::exception_ptr>> failed;
  long nl = 0;
  // one launch per kind, kinds in order of first submission
  std::vector<char> done(ops.size(), 0);
  std::vector<void*> same;
  for (size_t i = 0; i < ops.size(); ++i) {
    if (done[i] || !ops[i].first) continue;
    same.clear();
    for (size_t k = i; k < ops.size(); ++k)
      if (!done[k] && ops[k].first == ops[i].first) {
        same.push_back(ops[k].second);
        done[k] = 1;
      }
    try {
      nl += ops[i].first(same.data(), (int)same.size(), ctx_);
    } catch (...) {
      for (void* b : same) failed.emplace_back(b, std::current_exception());
    }
  }
  try {
    if (round_end_) round_end_(ctx_);
  } catch (...) {
    for (auto& o : ops) failed.emplace_back(o.second, std::current_exception());
  }
  lk.lock();
  t_end_ = now_s();
  launch_s_ += t_end_ - t0;
  failed_.swap(failed);
  arrived_ = 0;
  ++round_;
  ++rounds_;
  launches_ += nl;
  cv_.notify_all();
}

}  // namespace aifmm
