// Serving runtime: a request loop over resident factorisations.
//
// Requests are factorisations (a problem's parameters) and solves (right-hand
// sides against a factorised problem).  Worker slots take requests as they
// free up; each runs its work on the problem's own non-blocking HIP stream, so
// requests for different problems overlap on the device.  Solves queued for
// the same problem with the same tolerance are merged into one multi-RHS
// solve.  Factorisations are admitted by device memory: an estimate is
// reserved against the budget, least-recently used idle problems are evicted
// to make room, and the reservation is replaced by the bytes the problem
// actually holds once it is built.  A solve for an evicted problem factorises
// it again first.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
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
#include "prog.h"

namespace aifmm {
void set_last_error(const std::string& msg);  // capi.cpp
}

namespace {

double now_s() {
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
  int steps = 0, batch_rhs = 0, refactored = 0;
  double relres = -1;
};

struct Entry {
  aifmm_ls_params p{};
  std::vector<double> bumps;
  int precision = 1;
  int64_t N = 0;
  aifmm_ls_problem* h = nullptr;  // null: not resident
  int64_t bytes = 0;              // device bytes while resident (or reserved)
  int64_t last_bytes = 0;         // measured when last resident
  double last_use = 0;
  bool busy = false;
  bool dropped = false;
};

// admission estimate of a problem shape no problem has been built for yet:
// arena minimums plus a per-unknown figure (operator, factor, sweep and
// refinement vectors)
int64_t estimate_bytes(int64_t N, int precision) {
  return (int64_t(128) << 20) + N * (precision ? int64_t(40) << 10 : int64_t(72) << 10);
}

}  // namespace

struct aifmm_ls_server {
  std::mutex mu;
  std::condition_variable cv;       // queue or cache changed
  std::condition_variable cv_done;  // a request finished
  std::deque<std::shared_ptr<Req>> queue;
  std::unordered_map<int64_t, std::shared_ptr<Req>> reqs;
  std::map<int64_t, Entry> cache;
  int64_t budget = 0, resident = 0;
  int64_t next_ticket = 1;
  int max_batch_rhs = 64;
  bool stop = false;
  int running = 0, peak_running = 0;
  int64_t evictions = 0, refactors = 0, completed = 0, batches = 0, merged = 0;
  std::vector<std::thread> workers;
  // largest measured footprint per (levels, precision): the admission
  // estimate of problems of that shape
  std::map<std::pair<int, int>, int64_t> seen_bytes;

  // Bytes to reserve before building entry e.  Caller holds mu.
  int64_t admission_bytes(const Entry& e) const {
    if (e.last_bytes) return e.last_bytes;
    auto it = seen_bytes.find({e.p.n_levels, e.precision});
    if (it != seen_bytes.end()) return it->second;
    return std::min(estimate_bytes(e.N, e.precision), budget);
  }

  struct Work {
    int kind = -1;  // 0 factorise, 1 solve batch (factorising first if not resident)
    int64_t problem = 0;
    std::vector<std::shared_ptr<Req>> reqs;
  };

  // Evict idle resident problems, least recently used first, until `need`
  // more bytes fit.  Caller holds mu.
  bool make_room(int64_t need) {
    while (resident + need > budget) {
      Entry* lru = nullptr;
      for (auto& kv : cache) {
        Entry& e = kv.second;
        if (e.h && !e.busy && (!lru || e.last_use < lru->last_use)) lru = &e;
      }
      if (!lru) return false;
      aifmm_ls_destroy(lru->h);
      lru->h = nullptr;
      resident -= lru->bytes;
      lru->bytes = 0;
      ++evictions;
    }
    return true;
  }

  void fail(const std::shared_ptr<Req>& r, const std::string& why) {
    r->status = -1;
    r->err = why;
    r->t_start = r->t_end = now_s();
  }

  // Choose the next runnable work in arrival order.  Caller holds mu.
  bool pick(Work& w) {
    for (auto it = queue.begin(); it != queue.end(); ++it) {
      std::shared_ptr<Req> r = *it;
      auto ce = cache.find(r->problem);
      if (ce == cache.end() || ce->second.dropped) {
        fail(r, "unknown or released problem");
        queue.erase(it);
        ++completed;
        cv_done.notify_all();
        return pick(w);
      }
      Entry& e = ce->second;
      if (e.busy) continue;
      if (!e.h) {
        const int64_t need = admission_bytes(e);
        if (need > budget) {
          fail(r, "problem needs " + std::to_string(need) + " device bytes, more than the budget");
          queue.erase(it);
          ++completed;
          cv_done.notify_all();
          return pick(w);
        }
        if (!make_room(need)) continue;  // wait until memory frees up
        e.bytes = need;                  // reservation until measured
        resident += need;
      }
      w = Work();
      w.problem = r->problem;
      e.busy = true;
      if (r->kind == 0) {
        w.kind = 0;
        w.reqs.push_back(r);
        queue.erase(it);
        return true;
      }
      // a solve: merge the queued solves of this problem with equal settings
      w.kind = 1;
      int total = 0;
      for (auto jt = queue.begin(); jt != queue.end();) {
        const std::shared_ptr<Req>& q = *jt;
        if (q->kind == 1 && q->problem == r->problem && q->tol == r->tol && q->max_refine == r->max_refine &&
            (w.reqs.empty() || total + q->nrhs <= max_batch_rhs)) {
          total += q->nrhs;
          w.reqs.push_back(q);
          jt = queue.erase(jt);
        } else {
          ++jt;
        }
      }
      return true;
    }
    return false;
  }

  void run_factor(Entry& e, aifmm_ls_problem*& h, std::string& err) {
    aifmm_ls_params p = e.p;
    p.bumps = e.bumps.empty() ? nullptr : e.bumps.data();
    p.n_bumps = (int)(e.bumps.size() / 4);
    h = nullptr;
    if (aifmm_ls_create(&p, &h) || aifmm_ls_assemble(h) || aifmm_ls_factor(h, e.precision)) {
      err = aifmm_ls_last_error();
      if (h) aifmm_ls_destroy(h);
      h = nullptr;
    }
  }

  void execute(Work& w) {
    Entry* e;
    aifmm_ls_problem* h;
    {
      std::lock_guard<std::mutex> g(mu);
      e = &cache[w.problem];
      h = e->h;
    }
    const double t0 = now_s();
    for (auto& r : w.reqs) r->t_start = t0;
    std::string err;
    bool refactored = false;
    if (!h) {
      run_factor(*e, h, err);
      refactored = w.kind == 1;
    }
    if (h && w.kind == 1) {
      int k = 0;
      for (auto& r : w.reqs) k += r->nrhs;
      const int64_t N = e->N;
      std::vector<double> B((size_t)2 * N * k), X((size_t)2 * N * k);
      size_t off = 0;
      for (auto& r : w.reqs) {
        std::copy(r->b.begin(), r->b.end(), B.begin() + off);
        off += r->b.size();
      }
      double st[8] = {0};
      if (aifmm_ls_solve(h, k, B.data(), X.data(), w.reqs[0]->tol, w.reqs[0]->max_refine, st)) {
        err = aifmm_ls_last_error();
      } else {
        off = 0;
        for (auto& r : w.reqs) {
          std::copy(X.begin() + off, X.begin() + off + r->b.size(), r->x);
          off += r->b.size();
          r->steps = (int)st[1];
          r->relres = st[2];
          r->batch_rhs = k;
          r->refactored = refactored;
        }
      }
    }
    int64_t bytes = 0;
    if (h && aifmm_ls_device_bytes(h, &bytes)) bytes = 0;
    const double t1 = now_s();
    std::lock_guard<std::mutex> g(mu);
    if (h && e->dropped) {
      aifmm_ls_destroy(h);
      h = nullptr;
    }
    resident -= e->bytes;
    e->h = h;
    e->bytes = h ? bytes : 0;
    if (h) {
      e->last_bytes = bytes;
      int64_t& sb = seen_bytes[{e->p.n_levels, e->precision}];
      sb = std::max(sb, bytes);
    }
    resident += e->bytes;
    e->busy = false;
    e->last_use = t1;
    // a solve can grow a problem (vectors for more right-hand sides): give
    // the budget back by evicting idle problems
    make_room(0);
    refactors += refactored;
    ++batches;
    merged += (int64_t)w.reqs.size() - 1;
    for (auto& r : w.reqs) {
      r->t_end = t1;
      if (!err.empty()) {
        r->status = -1;
        r->err = err;
      } else {
        r->status = 0;
      }
      ++completed;
    }
    --running;
    cv.notify_all();
    cv_done.notify_all();
  }

  void worker() {
    for (;;) {
      Work w;
      {
        std::unique_lock<std::mutex> lk(mu);
        for (;;) {
          if (pick(w)) break;
          if (stop && queue.empty()) return;
          cv.wait(lk);
        }
        ++running;
        peak_running = std::max(peak_running, running);
      }
      execute(w);
    }
  }
};

namespace {
std::atomic<int> live_servers{0};
}

extern "C" {

#define SGUARD(...)                     \
  try {                                 \
    __VA_ARGS__;                        \
    return 0;                           \
  } catch (const std::exception& ex) { \
    aifmm::set_last_error(ex.what());   \
    return -1;                          \
  } catch (...) {                       \
    aifmm::set_last_error("unknown error");\
    return -1;                          \
  }

int aifmm_ls_server_create(int64_t budget_bytes, int slots, aifmm_ls_server** out) {
  SGUARD({
    if (!out) throw std::runtime_error("null argument");
    if (budget_bytes <= 0) throw std::runtime_error("device memory budget must be positive");
    if (slots < 1 || slots > 16) throw std::runtime_error("slots must be in [1, 16]");
    auto* s = new aifmm_ls_server();
    s->budget = budget_bytes;
    // the slots' threads must not capture graphs while another thread frees
    // memory or synchronises the device: sweeps launch directly while a
    // server exists
    if (live_servers++ == 0) aifmm::prog_graphs_hold(true);
    for (int i = 0; i < slots; ++i) s->workers.emplace_back([s] { s->worker(); });
    *out = s;
  })
}

int aifmm_ls_server_factor(aifmm_ls_server* s, const aifmm_ls_params* p, int precision, int64_t* problem,
                           int64_t* ticket) {
  SGUARD({
    if (!s || !p || !problem) throw std::runtime_error("null argument");
    if (precision != 0 && precision != 1) throw std::runtime_error("precision must be 0 (complex128) or 1 (complex64)");
    if (p->n_levels < 2 || p->n_levels > 9) throw std::runtime_error("n_levels must be in [2, 9]");
    auto r = std::make_shared<Req>();
    std::lock_guard<std::mutex> g(s->mu);
    if (s->stop) throw std::runtime_error("server is shutting down");
    r->ticket = s->next_ticket++;
    r->kind = 0;
    r->problem = r->ticket;
    r->t_submit = now_s();
    Entry& e = s->cache[r->problem];
    e.p = *p;
    if (p->bumps && p->n_bumps > 0) e.bumps.assign(p->bumps, p->bumps + 4 * p->n_bumps);
    e.p.bumps = nullptr;
    e.p.n_bumps = 0;
    e.precision = precision;
    e.N = (int64_t(1) << (2 * p->n_levels)) * p->n_cheb * p->n_cheb;
    s->reqs[r->ticket] = r;
    s->queue.push_back(r);
    *problem = r->problem;
    if (ticket) *ticket = r->ticket;
    s->cv.notify_all();
  })
}

int aifmm_ls_server_solve(aifmm_ls_server* s, int64_t problem, int nrhs, const double* b, double* x, double tol,
                          int max_refine, int64_t* ticket) {
  SGUARD({
    if (!s || !b || !x || !ticket) throw std::runtime_error("null argument");
    if (nrhs < 1) throw std::runtime_error("nrhs must be positive");
    auto r = std::make_shared<Req>();
    std::lock_guard<std::mutex> g(s->mu);
    if (s->stop) throw std::runtime_error("server is shutting down");
    auto ce = s->cache.find(problem);
    if (ce == s->cache.end() || ce->second.dropped) throw std::runtime_error("unknown problem");
    if (nrhs > s->max_batch_rhs) throw std::runtime_error("nrhs above the batch limit");
    r->ticket = s->next_ticket++;
    r->kind = 1;
    r->problem = problem;
    r->nrhs = nrhs;
    r->b.assign(b, b + 2 * ce->second.N * nrhs);
    r->x = x;
    r->tol = tol;
    r->max_refine = max_refine;
    r->t_submit = now_s();
    s->reqs[r->ticket] = r;
    s->queue.push_back(r);
    *ticket = r->ticket;
    s->cv.notify_all();
  })
}

int aifmm_ls_server_wait(aifmm_ls_server* s, int64_t ticket, double* info) {
  std::shared_ptr<Req> r;
  {
    std::unique_lock<std::mutex> lk(s->mu);
    auto it = s->reqs.find(ticket);
    if (it == s->reqs.end()) {
      aifmm::set_last_error("unknown ticket");
      return -1;
    }
    r = it->second;
    s->cv_done.wait(lk, [&] { return r->status != 1; });
    s->reqs.erase(ticket);
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
  }
  if (r->status < 0) {
    aifmm::set_last_error(r->err);
    return -1;
  }
  return 0;
}

int aifmm_ls_server_release(aifmm_ls_server* s, int64_t problem) {
  SGUARD({
    std::lock_guard<std::mutex> g(s->mu);
    auto ce = s->cache.find(problem);
    if (ce == s->cache.end()) throw std::runtime_error("unknown problem");
    Entry& e = ce->second;
    e.dropped = true;
    if (e.h && !e.busy) {
      aifmm_ls_destroy(e.h);
      e.h = nullptr;
      s->resident -= e.bytes;
      e.bytes = 0;
    }
    s->cv.notify_all();
  })
}

int aifmm_ls_server_stats(aifmm_ls_server* s, double* out, int n) {
  SGUARD({
    double v[10] = {0};
    std::lock_guard<std::mutex> g(s->mu);
    for (auto& kv : s->cache) v[0] += kv.second.h != nullptr;
    v[1] = (double)s->resident;
    v[2] = (double)s->budget;
    v[3] = (double)s->evictions;
    v[4] = (double)s->refactors;
    v[5] = (double)s->completed;
    v[6] = (double)s->batches;
    v[7] = (double)s->merged;
    v[8] = s->peak_running;
    v[9] = (double)s->queue.size();
    for (int i = 0; i < n && i < 10; ++i) out[i] = v[i];
  })
}

int aifmm_ls_server_destroy(aifmm_ls_server* s) {
  if (!s) return 0;
  {
    std::unique_lock<std::mutex> lk(s->mu);
    s->cv_done.wait(lk, [&] { return s->queue.empty() && s->running == 0; });
    s->stop = true;
    s->cv.notify_all();
  }
  for (auto& t : s->workers) t.join();
  for (auto& kv : s->cache)
    if (kv.second.h) aifmm_ls_destroy(kv.second.h);
  delete s;
  if (--live_servers == 0) aifmm::prog_graphs_hold(false);
  return 0;
}

}  // extern "C"
