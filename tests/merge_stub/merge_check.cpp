// Host-only test of the lockstep merge of batched launches (merge.h / merge.cpp,
// built with g++ by tests/test_merge_core.py; no HIP).
//
// Member threads submit seeded sequences of batches of three kinds, some of
// them empty, and some members stop early or fail.  A stub launch records, per
// round, which batches of which kind it took together.  Checked: every batch is
// launched exactly once; a round launches each kind at most once; a member's
// batches are launched in its own order, one per round; the members whose
// batches were in a failed launch see its error, and the rest of the group
// goes on.
#include <atomic>
#include <chrono>
#include <cstdint>
#include <map>
#include <mutex>
#include <random>
#include <stdexcept>
#include <thread>
#include <vector>

#include "../../aifmm_ls_amd/csrc/merge.h"

using aifmm::MergeGroup;
using aifmm::MergeMember;

namespace {

struct Batch {
  int member, seq, kind;
  bool empty, fail;
  long launched_round = -1;
};

struct Log {
  std::mutex mu;
  long round = 0;  // advanced by round_end
  std::map<std::pair<long, int>, int> kind_launches;  // (round, kind) -> launches
  long launched = 0, empty_seen = 0;
} lg;

template <int K>
int stub_launch(void* const* bs, int n, void* ctx) {
  (void)ctx;
  std::lock_guard<std::mutex> g(lg.mu);
  bool any = false, fail = false;
  for (int i = 0; i < n; ++i) {
    Batch* b = (Batch*)bs[i];
    if (b->kind != K) throw std::logic_error("batch of another kind in this launch");
    if (b->launched_round >= 0) throw std::logic_error("batch launched twice");
    b->launched_round = lg.round;
    if (b->empty) ++lg.empty_seen;
    else any = true;
    fail |= b->fail;
  }
  if (!any) return 0;
  ++lg.kind_launches[{lg.round, K}];
  ++lg.launched;
  if (fail) throw std::runtime_error("stub launch failure");
  // a little device-like latency so members overlap
  std::this_thread::sleep_for(std::chrono::microseconds(50));
  return 1;
}
void round_end(void*) {
  std::lock_guard<std::mutex> g(lg.mu);
  ++lg.round;
}
const aifmm::MergeFn KINDS[3] = {&stub_launch<0>, &stub_launch<1>, &stub_launch<2>};

}  // namespace

extern "C" int merge_check(int seed, int members, int max_ops, double* out) {
  lg.round = 0;
  lg.kind_launches.clear();
  lg.launched = lg.empty_seen = 0;
  std::vector<std::vector<Batch>> plan(members);
  std::mt19937_64 rng(seed);
  for (int m = 0; m < members; ++m) {
    int nops = 1 + (int)(rng() % max_ops);
    if (m % 5 == 4) nops = 1 + nops / 4;  // members that stop early
    for (int q = 0; q < nops; ++q) {
      Batch b{m, q, (int)(rng() % 3), rng() % 4 == 0, false, -1};
      b.fail = (m == 2 && q == nops / 2 && members > 3);  // one failing launch
      if (b.fail) b.empty = false;
      plan[m].push_back(b);
    }
  }
  int ctx = 0;
  MergeGroup grp(members, &ctx, &round_end);
  std::vector<int> threw(members, 0), done_ops(members, 0);
  std::vector<std::thread> th;
  for (int m = 0; m < members; ++m)
    th.emplace_back([&, m] {
      MergeMember mm(&grp);
      std::mt19937 trng(seed * 1000 + m);
      try {
        for (auto& b : plan[m]) {
          std::this_thread::sleep_for(std::chrono::microseconds(trng() % 3 == 0 ? 30 : 0));
          grp.submit(KINDS[b.kind], &b);
          ++done_ops[m];
        }
      } catch (const std::runtime_error&) {
        threw[m] = 1;
      }
    });
  for (auto& t : th) t.join();
  // checks
  long violations = 0, failing_round = -1;
  int failing_kind = -1;
  for (auto& p : plan)
    for (auto& b : p)
      if (b.fail) failing_round = b.launched_round, failing_kind = b.kind;
  for (int m = 0; m < members; ++m) {
    long prev = -1;
    for (int q = 0; q < done_ops[m] + threw[m] && q < (int)plan[m].size(); ++q) {
      const Batch& b = plan[m][q];
      if (b.launched_round < 0) ++violations;          // submitted but never launched
      if (b.launched_round <= prev) ++violations;      // out of order or two per round
      prev = b.launched_round;
    }
    for (int q = done_ops[m] + threw[m]; q < (int)plan[m].size(); ++q)
      if (plan[m][q].launched_round >= 0) ++violations;  // launched after its member failed
    // a member sees an error exactly when its batch was in the failed launch
    bool in_fail = false;
    for (auto& b : plan[m])
      if (b.launched_round >= 0 && b.launched_round == failing_round && b.kind == failing_kind) in_fail = true;
    if (in_fail != (threw[m] != 0)) ++violations;
    if (!threw[m] && done_ops[m] != (int)plan[m].size()) ++violations;
  }
  for (auto& kv : lg.kind_launches)
    if (kv.second > 1) ++violations;
  long total = 0;
  for (int m = 0; m < members; ++m) total += done_ops[m] + threw[m];
  out[0] = (double)violations;
  out[1] = (double)grp.rounds();
  out[2] = (double)grp.launches();
  out[3] = (double)lg.launched;
  out[4] = (double)total;
  out[5] = (double)failing_round;
  int nthrew = 0;
  for (int v : threw) nthrew += v;
  out[6] = nthrew;
  return 0;
}
