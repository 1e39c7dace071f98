// Lockstep merging of the batched launches of several host threads.
//
// A merged factorisation runs one host thread per problem, all on one
// stream.  Every batched launch a member thread makes (a GEMM batch, a copy
// batch, an LU batch, ...) becomes a submission to the group; once every
// active member has submitted, the last one to arrive launches the
// submissions of each kind as one launch over all their descriptors, and the
// members continue.  A member's launches keep their order, and its own stream
// operations (uploads, memsets, read-backs) sit between them on the shared
// stream; members never wait on each other's device work, since only launched
// work is ever queued.  Problems of equal structure stay in step, so each of
// their launches carries the descriptors of all of them.
#pragma once
#include <condition_variable>
#include <exception>
#include <mutex>
#include <utility>
#include <vector>

namespace aifmm {

// Launch n batches of one kind together (each batch is empty afterwards);
// ctx: the group's launch context.  Returns the number of kernel launches.
using MergeFn = int (*)(void* const* batches, int n, void* ctx);

class MergeGroup {
 public:
  // ctx is handed to every MergeFn; round_end(ctx), if given, runs after each
  // round's launches (by the member that launched them)
  MergeGroup(int members, void* ctx, void (*round_end)(void*) = nullptr)
      : active_(members), ctx_(ctx), round_end_(round_end) {}
  MergeGroup(const MergeGroup&) = delete;
  MergeGroup& operator=(const MergeGroup&) = delete;
  // Hand in a batch (fn: its kind's merged launch; a null fn or an empty
  // batch only keeps this member in step).  Returns once the round has been
  // launched; rethrows the error of a failed launch this batch was part of
  // (the round's other launches still go out).
  void submit(MergeFn fn, void* batch);
  // A member that has finished (or failed) stops taking part.
  void leave();
  long rounds() const { return rounds_; }
  long launches() const { return launches_; }  // kernel launches of the merged batches
  // seconds spent launching the rounds, and from the end of one round to the
  // last submission of the next (the members' host work between launches)
  double launch_s() const { return launch_s_; }
  double gather_s() const { return gather_s_; }

 private:
  void execute(std::unique_lock<std::mutex>& lk);
  std::mutex mu_;
  std::condition_variable cv_;
  int active_, arrived_ = 0;
  unsigned long round_ = 0;
  std::vector<std::pair<MergeFn, void*>> ops_;
  std::vector<std::pair<void*, std::exception_ptr>> failed_;  // batches of failed launches, this round
  void* ctx_;
  void (*round_end_)(void*);
  long rounds_ = 0, launches_ = 0;
  double launch_s_ = 0, gather_s_ = 0, t_end_ = -1;
};

// The calling thread's group (null outside a merged run).
MergeGroup* merge_group();

// Membership of the calling thread for the scope's lifetime; leaving the
// scope (also by an exception) leaves the group.
class MergeMember {
 public:
  explicit MergeMember(MergeGroup* g);
  ~MergeMember();
  MergeMember(const MergeMember&) = delete;
  MergeMember& operator=(const MergeMember&) = delete;

 private:
  MergeGroup* g_;
};

}  // namespace aifmm
