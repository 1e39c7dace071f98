// Scheduling core of the serving loop (no reference counterpart: the
// reference factorises and solves one problem per process,
// testFMM2D_find_error.cpp:41-176).  Header-only and free of HIP: the device
// work goes through an Ops policy, the C ABI in serve.cpp in the library and
// stub handles in the host-only scheduler test.
//
// Requests are factorisations (a problem's parameters) and solves (right-hand
// sides against a factorised problem).  Worker slots take work as they free
// up; each runs on its problem's own stream, so work for different problems
// overlaps on the device.  Queued solves of one problem with equal settings
// merge into one multi-right-hand-side solve; queued one-column solves of
// different problems with equal settings and precision merge into one batch
// solve (one launch sequence for all of them).  Problems are admitted by device
// memory: an estimate is reserved against the budget, least-recently-used idle
// problems are evicted to make room, and the reservation becomes the measured
// bytes once built.  A solve for an evicted problem factorises it again.
//
// Locking: one mutex guards the queue, the cache and the byte counts.  Device
// work, handle destruction (device frees) and host copies all run with the
// mutex released.  An evicted or released handle is moved to `doomed`; its
// bytes stay counted as resident until a thread without the lock has destroyed
// it, so the budget is never exceeded while frees are pending.
#pragma once
#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstdint>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "../../include/aifmm_ls.h"

namespace aifmm {
namespace serving {

inline double now_s() {
  using namespace std::chrono;
  return duration<double>(steady_clock::now().time_since_epoch()).count();
}

struct Req {
  int64_t ticket = 0;
  int kind = 0;  // 0: factorisation, 1: solve
  int64_t problem = 0;
  int nrhs = 0;
  std::vector<double> b;  // 2 N nrhs
  double* x = nullptr;
  double tol = 0;
  int max_refine = 0;
  double t_submit = 0, t_start = 0, t_end = 0;
  int status = 1;  // 1: queued or running, 0: done, -1: failed
  std::string err;
  int steps = 0, batch_rhs = 0, refactored = 0, batch_problems = 0;
  double relres = -1;
};

// admission estimate of a problem shape no problem has been built for yet:
// arena minimums plus a per-unknown figure (operator, factor, sweep and
// refinement vectors)
inline int64_t estimate_bytes(int64_t N, int precision) {
  return (int64_t(128) << 20) + N * (precision ? int64_t(40) << 10 : int64_t(72) << 10);
}

struct Counters {
  int64_t evictions = 0, refactors = 0, completed = 0, batches = 0, merged = 0, cross_batches = 0;
  int peak_running = 0;
};

// Ops: static functions over an opaque Handle (see serve.cpp):
//   int build(const aifmm_ls_params&, int precision, Handle&, std::string& err)   (0 = ok)
//   int solve(Handle, int nrhs, const double* b, double* x, double tol, int max_refine,
//             double* st8, std::string& err)
//   int solve_batch(Handle* hs, int n, const double* const* b, double* const* x, double tol,
//                   int max_refine, double* relres, double* st3, std::string& err)
//   int64_t bytes(Handle)            device bytes the handle holds
//   void destroy(Handle)
template <class Ops>
class Server {
 public:
  using Handle = typename Ops::Handle;

  Server(int64_t budget, int slots, int max_batch_rhs = 64, int max_batch_problems = 32)
      : budget_(budget), max_batch_rhs_(max_batch_rhs), max_batch_problems_(max_batch_problems) {
    max_builds_ = slots / 2 > 1 ? slots / 2 : 1;
    for (int i = 0; i < slots; ++i) workers_.emplace_back([this] { worker(); });
  }

  // Finish every queued request, stop the slots, free all problems.
  ~Server() {
    {
      std::unique_lock<std::mutex> lk(mu_);
      cv_done_.wait(lk, [&] { return queue_.empty() && running_ == 0; });
      stop_ = true;
      cv_.notify_all();
    }
    for (auto& t : workers_) t.join();
    for (auto& d : doomed_) Ops::destroy(d.h);
    for (auto& kv : cache_)
      if (kv.second.h) Ops::destroy(kv.second.h);
  }

  int64_t submit_factor(const aifmm_ls_params& p, int precision, int64_t* ticket) {
    auto r = std::make_shared<Req>();
    std::lock_guard<std::mutex> g(mu_);
    if (stop_) throw std::runtime_error("server is shutting down");
    r->ticket = next_ticket_++;
    r->kind = 0;
    r->problem = r->ticket;
    r->t_submit = now_s();
    Entry& e = cache_[r->problem];
    e.p = p;
    if (p.bumps && p.n_bumps > 0) e.bumps.assign(p.bumps, p.bumps + 4 * p.n_bumps);
    e.p.bumps = nullptr;
    e.p.n_bumps = 0;
    e.precision = precision;
    e.N = (int64_t(1) << (2 * p.n_levels)) * p.n_cheb * p.n_cheb;
    reqs_[r->ticket] = r;
    queue_.push_back(r);
    if (ticket) *ticket = r->ticket;
    cv_.notify_all();
    return r->problem;
  }

  int64_t submit_solve(int64_t problem, int nrhs, const double* b, double* x, double tol, int max_refine) {
    if (nrhs < 1) throw std::runtime_error("nrhs must be positive");
    auto r = std::make_shared<Req>();
    int64_t N;
    {
      std::lock_guard<std::mutex> g(mu_);
      auto ce = cache_.find(problem);
      if (ce == cache_.end() || ce->second.dropped) throw std::runtime_error("unknown problem");
      if (nrhs > max_batch_rhs_) throw std::runtime_error("nrhs above the batch limit");
      N = ce->second.N;
    }
    // the copy of the right-hand sides runs outside the lock
    r->b.assign(b, b + 2 * N * nrhs);
    std::lock_guard<std::mutex> g(mu_);
    if (stop_) throw std::runtime_error("server is shutting down");
    auto ce = cache_.find(problem);
    if (ce == cache_.end() || ce->second.dropped) throw std::runtime_error("unknown problem");
    r->ticket = next_ticket_++;
    r->kind = 1;
    r->problem = problem;
    r->nrhs = nrhs;
    r->x = x;
    r->tol = tol;
    r->max_refine = max_refine;
    r->t_submit = now_s();
    reqs_[r->ticket] = r;
    queue_.push_back(r);
    cv_.notify_all();
    return r->ticket;
  }

  // Block until the request ends; info as aifmm_ls_server_wait.
  int wait(int64_t ticket, double* info, std::string& err) {
    std::shared_ptr<Req> r;
    {
      std::unique_lock<std::mutex> lk(mu_);
      auto it = reqs_.find(ticket);
      if (it == reqs_.end()) {
        err = "unknown ticket";
        return -1;
      }
      r = it->second;
      cv_done_.wait(lk, [&] { return r->status != 1; });
      reqs_.erase(ticket);
    }
    if (info) {
      info[0] = r->status;
      info[1] = r->t_end - r->t_submit;
      info[2] = r->t_start - r->t_submit;
      info[3] = r->t_end - r->t_start;
      info[4] = r->steps;
      info[5] = r->relres;
      info[6] = r->batch_rhs;
      info[7] = r->refactored;
      info[8] = r->batch_problems;
    }
    if (r->status < 0) {
      err = r->err;
      return -1;
    }
    return 0;
  }

  // Drop a problem: freed now (outside the lock), or when its work ends.
  void release(int64_t problem) {
    std::unique_lock<std::mutex> lk(mu_);
    auto ce = cache_.find(problem);
    if (ce == cache_.end()) throw std::runtime_error("unknown problem");
    Entry& e = ce->second;
    e.dropped = true;
    if (e.h && !e.busy) doom(e);
    reap(lk);
    cv_.notify_all();
  }

  // resident problems, resident bytes, budget, evictions, re-factorisations,
  // completed, executed batches, merged requests, peak running, queued,
  // cross-problem batches
  void stats(double* v) {
    std::lock_guard<std::mutex> g(mu_);
    v[0] = 0;
    for (auto& kv : cache_) v[0] += kv.second.h != Handle();
    v[1] = (double)resident_;
    v[2] = (double)budget_;
    v[3] = (double)c_.evictions;
    v[4] = (double)c_.refactors;
    v[5] = (double)c_.completed;
    v[6] = (double)c_.batches;
    v[7] = (double)c_.merged;
    v[8] = c_.peak_running;
    v[9] = (double)queue_.size();
    v[10] = (double)c_.cross_batches;
  }

 private:
  struct Entry {
    aifmm_ls_params p{};
    std::vector<double> bumps;
    int precision = 1;
    int64_t N = 0;
    Handle h = Handle();    // none: not resident
    int64_t bytes = 0;      // device bytes while resident (or reserved)
    int64_t last_bytes = 0; // measured when last resident
    double last_use = 0;
    bool busy = false;
    bool dropped = false;
  };
  struct Doomed {
    Handle h;
    int64_t bytes;
  };
  struct Work {
    int kind = -1;  // 0 factorise, 1 solve one problem (factorising first if not resident), 2 batch of problems
    bool build = false;  // a factorisation runs first (new or evicted problem)
    std::vector<int64_t> problems;
    std::vector<std::shared_ptr<Req>> reqs;  // kind 2: one per problem, same order
  };

  std::mutex mu_;
  std::condition_variable cv_;       // queue, cache or bytes changed
  std::condition_variable cv_done_;  // a request finished
  std::deque<std::shared_ptr<Req>> queue_;
  std::unordered_map<int64_t, std::shared_ptr<Req>> reqs_;
  std::map<int64_t, Entry> cache_;
  std::vector<Doomed> doomed_;
  int64_t budget_ = 0, resident_ = 0, doomed_bytes_ = 0;
  int64_t next_ticket_ = 1;
  int max_batch_rhs_, max_batch_problems_;
  bool stop_ = false;
  int running_ = 0;
  // Work that factorises (a new problem, or a solve of an evicted one) holds
  // a slot for the length of a factorisation: at most max_builds_ slots do at
  // once (half the slots, at least one), so solves of resident problems keep
  // slots while factorisations queue.
  int builds_running_ = 0, max_builds_ = 1;
  Counters c_;
  std::vector<std::thread> workers_;
  // largest measured footprint per (levels, precision): the admission
  // estimate of problems of that shape
  std::map<std::pair<int, int>, int64_t> seen_bytes_;

  // Bytes to reserve before building entry e.  Caller holds mu_.
  int64_t admission_bytes(const Entry& e) const {
    if (e.last_bytes) return e.last_bytes;
    auto it = seen_bytes_.find({e.p.n_levels, e.precision});
    if (it != seen_bytes_.end()) return it->second;
    return std::min(estimate_bytes(e.N, e.precision), budget_);
  }

  // Move a resident idle problem's handle to the destroy list; its bytes stay
  // counted until it has been destroyed.  Caller holds mu_.
  void doom(Entry& e) {
    doomed_.push_back(Doomed{e.h, e.bytes});
    doomed_bytes_ += e.bytes;
    e.h = Handle();
    e.bytes = 0;
  }

  // Destroy the doomed handles with mu_ released, then give their bytes back.
  void reap(std::unique_lock<std::mutex>& lk) {
    while (!doomed_.empty()) {
      std::vector<Doomed> v;
      v.swap(doomed_);
      lk.unlock();
      for (auto& d : v) Ops::destroy(d.h);
      lk.lock();
      for (auto& d : v) {
        resident_ -= d.bytes;
        doomed_bytes_ -= d.bytes;
      }
      cv_.notify_all();
    }
  }

  // Doom idle resident problems, least recently used first, until `need`
  // more bytes fit once the doomed ones are freed; true when they fit now.
  // Caller holds mu_.
  bool make_room(int64_t need) {
    while (resident_ - doomed_bytes_ + need > budget_) {
      Entry* lru = nullptr;
      for (auto& kv : cache_) {
        Entry& e = kv.second;
        if (e.h != Handle() && !e.busy && (!lru || e.last_use < lru->last_use)) lru = &e;
      }
      if (!lru) return false;
      doom(*lru);
      ++c_.evictions;
    }
    return resident_ + need <= budget_;
  }

  void fail(const std::shared_ptr<Req>& r, const std::string& why) {
    r->status = -1;
    r->err = why;
    r->t_start = r->t_end = now_s();
    ++c_.completed;
    cv_done_.notify_all();
  }

  // Reserve memory for a non-resident entry; false: not now.  Caller holds mu_.
  bool admit(Entry& e) {
    if (e.h != Handle()) return true;
    const int64_t need = admission_bytes(e);
    if (!make_room(need)) return false;
    e.bytes = need;  // reservation until measured
    resident_ += need;
    return true;
  }

  // Choose the next runnable work in arrival order; work that factorises is
  // passed over while max_builds_ factorisations run.  Caller holds mu_.
  bool pick(Work& w) {
    for (auto it = queue_.begin(); it != queue_.end();) {
      std::shared_ptr<Req> r = *it;
      auto ce = cache_.find(r->problem);
      if (ce == cache_.end() || ce->second.dropped) {
        fail(r, "unknown or released problem");
        it = queue_.erase(it);
        continue;
      }
      Entry& e = ce->second;
      if (e.busy) {
        ++it;
        continue;
      }
      const bool needs_build = e.h == Handle();
      if (needs_build && builds_running_ >= max_builds_) {
        ++it;
        continue;
      }
      if (e.h == Handle() && admission_bytes(e) > budget_) {
        fail(r, "problem needs " + std::to_string(admission_bytes(e)) + " device bytes, more than the budget");
        it = queue_.erase(it);
        continue;
      }
      if (!admit(e)) {
        ++it;  // wait until memory frees up
        continue;
      }
      w = Work();
      w.problems.push_back(r->problem);
      e.busy = true;
      w.build = needs_build;
      builds_running_ += needs_build;
      if (r->kind == 0) {
        w.kind = 0;
        w.reqs.push_back(r);
        queue_.erase(it);
        return true;
      }
      // a solve: the queued solves of this problem with equal settings
      int total = 0;
      for (auto jt = queue_.begin(); jt != queue_.end();) {
        const std::shared_ptr<Req>& q = *jt;
        if (q->kind == 1 && q->problem == r->problem && q->tol == r->tol && q->max_refine == r->max_refine &&
            (w.reqs.empty() || total + q->nrhs <= max_batch_rhs_)) {
          total += q->nrhs;
          w.reqs.push_back(q);
          jt = queue_.erase(jt);
        } else {
          ++jt;
        }
      }
      w.kind = 1;
      if (total == 1 && e.h != Handle()) gather_batch(w, e.precision);
      return true;
    }
    return false;
  }

  // A one-column solve of a resident problem: add queued one-column solves of
  // other idle resident problems with the same settings and precision, one
  // per problem, as one batch.  Caller holds mu_.
  void gather_batch(Work& w, int precision) {
    const std::shared_ptr<Req> r0 = w.reqs[0];
    for (auto jt = queue_.begin(); jt != queue_.end() && (int)w.problems.size() < max_batch_problems_;) {
      const std::shared_ptr<Req>& q = *jt;
      auto ce = cache_.find(q->problem);
      const bool ok = q->kind == 1 && q->nrhs == 1 && q->tol == r0->tol && q->max_refine == r0->max_refine &&
                      ce != cache_.end() && !ce->second.dropped && !ce->second.busy &&
                      ce->second.h != Handle() && ce->second.precision == precision &&
                      std::find(w.problems.begin(), w.problems.end(), q->problem) == w.problems.end();
      if (!ok) {
        ++jt;
        continue;
      }
      ce->second.busy = true;
      w.problems.push_back(q->problem);
      w.reqs.push_back(q);
      jt = queue_.erase(jt);
    }
    if (w.problems.size() > 1) w.kind = 2;
  }

  void build(Entry& e, Handle& h, std::string& err) {
    aifmm_ls_params p = e.p;
    p.bumps = e.bumps.empty() ? nullptr : e.bumps.data();
    p.n_bumps = (int)(e.bumps.size() / 4);
    h = Handle();
    if (Ops::build(p, e.precision, h, err)) h = Handle();
  }

  void execute(Work& w) {
    const size_t np = w.problems.size();
    std::vector<Entry*> es(np);
    std::vector<Handle> hs(np);
    {
      std::lock_guard<std::mutex> g(mu_);
      for (size_t i = 0; i < np; ++i) {
        es[i] = &cache_[w.problems[i]];
        hs[i] = es[i]->h;
      }
    }
    const double t0 = now_s();
    for (auto& r : w.reqs) r->t_start = t0;
    std::vector<std::string> err(np);
    bool refactored = false;
    if (hs[0] == Handle()) {  // kinds 0 and 1 only: batches hold resident problems
      build(*es[0], hs[0], err[0]);
      refactored = w.kind == 1;
    }
    if (w.kind == 1 && hs[0] != Handle()) {
      int k = 0;
      for (auto& r : w.reqs) k += r->nrhs;
      const int64_t N = es[0]->N;
      std::vector<double> B((size_t)2 * N * k), X((size_t)2 * N * k);
      size_t off = 0;
      for (auto& r : w.reqs) {
        std::copy(r->b.begin(), r->b.end(), B.begin() + off);
        off += r->b.size();
      }
      double st[8] = {0};
      if (!Ops::solve(hs[0], k, B.data(), X.data(), w.reqs[0]->tol, w.reqs[0]->max_refine, st, err[0])) {
        off = 0;
        for (auto& r : w.reqs) {
          std::copy(X.begin() + off, X.begin() + off + r->b.size(), r->x);
          off += r->b.size();
          r->steps = (int)st[1];
          r->relres = st[2];
          r->batch_rhs = k;
          r->batch_problems = 1;
          r->refactored = refactored;
        }
      }
    }
    if (w.kind == 2) {
      std::vector<const double*> b(np);
      std::vector<double*> x(np);
      for (size_t i = 0; i < np; ++i) b[i] = w.reqs[i]->b.data(), x[i] = w.reqs[i]->x;
      std::vector<double> rel(np, -1.0);
      double st[3] = {0};
      std::string e;
      if (Ops::solve_batch(hs.data(), (int)np, b.data(), x.data(), w.reqs[0]->tol, w.reqs[0]->max_refine,
                           rel.data(), st, e)) {
        for (auto& s : err) s = e;
      } else {
        for (size_t i = 0; i < np; ++i) {
          w.reqs[i]->steps = (int)st[1];
          w.reqs[i]->relres = rel[i];
          w.reqs[i]->batch_rhs = 1;
          w.reqs[i]->batch_problems = (int)np;
        }
      }
    }
    std::vector<int64_t> bytes(np, 0);
    for (size_t i = 0; i < np; ++i)
      if (hs[i] != Handle()) bytes[i] = Ops::bytes(hs[i]);
    const double t1 = now_s();
    std::unique_lock<std::mutex> lk(mu_);
    for (size_t i = 0; i < np; ++i) {
      Entry& e = *es[i];
      resident_ -= e.bytes;  // the reservation or the bytes before this work
      e.h = hs[i];
      e.bytes = hs[i] != Handle() ? bytes[i] : 0;
      resident_ += e.bytes;
      if (hs[i] != Handle()) {
        e.last_bytes = bytes[i];
        int64_t& sb = seen_bytes_[{e.p.n_levels, e.precision}];
        sb = std::max(sb, bytes[i]);
      }
      e.busy = false;
      e.last_use = t1;
      if (e.dropped && e.h != Handle()) doom(e);
    }
    // a solve can grow a problem (vectors for more right-hand sides): give
    // the budget back by evicting idle problems
    make_room(0);
    c_.refactors += refactored;
    ++c_.batches;
    c_.cross_batches += w.kind == 2;
    c_.merged += (int64_t)w.reqs.size() - 1;
    for (size_t q = 0; q < w.reqs.size(); ++q) {
      auto& r = w.reqs[q];
      const std::string& e = err[w.kind == 2 ? q : 0];
      r->t_end = t1;
      if (!e.empty()) {
        r->status = -1;
        r->err = e;
      } else {
        r->status = 0;
      }
      ++c_.completed;
    }
    --running_;
    builds_running_ -= w.build;
    reap(lk);
    cv_.notify_all();
    cv_done_.notify_all();
  }

  void worker() {
    std::unique_lock<std::mutex> lk(mu_);
    for (;;) {
      if (!doomed_.empty()) {
        reap(lk);
        continue;
      }
      Work w;
      if (pick(w)) {
        ++running_;
        c_.peak_running = std::max(c_.peak_running, running_);
        lk.unlock();
        execute(w);
        lk.lock();
        continue;
      }
      // pick() may have doomed idle problems to make room: free them before
      // waiting, or no thread would
      if (!doomed_.empty()) continue;
      if (stop_ && queue_.empty()) return;
      cv_.wait(lk);
    }
  }
};

}  // namespace serving
}  // namespace aifmm
