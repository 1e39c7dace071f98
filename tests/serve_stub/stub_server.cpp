// Host-only test of the serving loop's scheduling core (serve_core.h) with
// stub problem handles instead of GPU factorisations.  Built by
// tests/test_serve_core.py with g++; no HIP.
//
// A stub build "allocates" its bytes at once and holds them until destroy; a
// stub solve maps b to x = s b with a per-problem scale s; destroy sleeps, as
// device frees synchronise.  The driver replays a seeded trace of factor and
// solve requests from several submitting threads, with releases, against a
// budget that forces evictions, and reports what a scheduler must never do.
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstring>
#include <mutex>
#include <random>
#include <set>
#include <thread>
#include <vector>

#include "../../aifmm_ls_amd/csrc/serve_core.h"

namespace {

struct Cfg {
  int build_ms = 5, solve_ms = 2, destroy_ms = 30;
  int64_t bytes_per_problem = 100;
} cfg;

struct Stub {
  double scale;
  int64_t bytes;
  int precision;
  std::atomic<int> users{0};
};

std::mutex reg_mu;
std::set<Stub*> live;
std::atomic<int64_t> live_bytes{0}, peak_bytes{0};
std::atomic<int> violations{0}, builds{0}, destroys{0}, batch_calls{0};
std::atomic<int> building{0}, peak_building{0};  // builds running at once

void sleep_ms(int ms) { std::this_thread::sleep_for(std::chrono::milliseconds(ms)); }

bool alive(Stub* h) {
  std::lock_guard<std::mutex> g(reg_mu);
  return live.count(h) != 0;
}

struct StubOps {
  using Handle = Stub*;
  static int build(const aifmm_ls_params& p, int precision, Handle& h, std::string& err) {
    if (p.kappa < 0) {
      err = "stub build failure";
      return -1;
    }
    h = new Stub{p.kappa, cfg.bytes_per_problem, precision};
    {
      std::lock_guard<std::mutex> g(reg_mu);
      live.insert(h);
    }
    const int64_t now = live_bytes += h->bytes;
    int64_t pk = peak_bytes.load();
    while (now > pk && !peak_bytes.compare_exchange_weak(pk, now)) {
    }
    ++builds;
    const int nb = ++building;
    int pb = peak_building.load();
    while (nb > pb && !peak_building.compare_exchange_weak(pb, nb)) {
    }
    sleep_ms(cfg.build_ms);
    --building;
    return 0;
  }
  static int solve(Handle h, int nrhs, const double* b, double* x, double, int, double* st, std::string&) {
    if (!alive(h)) ++violations;
    if (h->users++ != 0) ++violations;  // two slots on one handle
    sleep_ms(cfg.solve_ms);
    // the driver tags every right-hand side block with N in its first double
    const int64_t n = (int64_t)std::llround(b[0]) * 2 * nrhs;
    for (int64_t i = 0; i < n; ++i) x[i] = b[i] * h->scale;
    st[1] = 1;
    st[2] = 1e-12;
    --h->users;
    return 0;
  }
  static int solve_batch(Handle* hs, int n, const double* const* b, double* const* x, double, int, double* relres,
                         double* st, std::string&) {
    ++batch_calls;
    std::set<Stub*> seen;
    for (int i = 0; i < n; ++i) {
      if (!alive(hs[i]) || !seen.insert(hs[i]).second) ++violations;
      if (hs[i]->users++ != 0) ++violations;
      if (hs[i]->precision != hs[0]->precision) ++violations;
    }
    sleep_ms(cfg.solve_ms);
    for (int i = 0; i < n; ++i) {
      // one column: N from the tag in the first double
      const int64_t N = (int64_t)std::llround(b[i][0]);
      for (int64_t k = 0; k < 2 * N; ++k) x[i][k] = b[i][k] * hs[i]->scale;
      relres[i] = 1e-12;
      --hs[i]->users;
    }
    st[1] = 1;
    return 0;
  }
  static int64_t bytes(Handle h) { return h->bytes; }
  static void destroy(Handle h) {
    if (!alive(h) || h->users.load() != 0) ++violations;
    sleep_ms(cfg.destroy_ms);
    {
      std::lock_guard<std::mutex> g(reg_mu);
      live.erase(h);
    }
    live_bytes -= h->bytes;
    ++destroys;
    delete h;
  }
};

}  // namespace

extern "C" {

// out: completed, failed, max submit latency (ms), peak live bytes, budget,
// evictions, violations, cross-problem batches, wrong results, builds,
// destroys, live bytes at the end, stub batch calls, peak concurrent builds
// (out[14]), solves refused because
// their problem was released after they were queued
int stub_trace(int seed, int nreq, int slots, int nsub, int build_ms, int solve_ms, int destroy_ms,
               int64_t budget_problems, double* out) {
  cfg.build_ms = build_ms;
  cfg.solve_ms = solve_ms;
  cfg.destroy_ms = destroy_ms;
  cfg.bytes_per_problem = int64_t(100) << 20;  // below the admission estimate of an unseen shape, as real problems are
  violations = 0;
  peak_bytes = 0;
  builds = destroys = batch_calls = 0;
  building = peak_building = 0;
  const int64_t budget = budget_problems * cfg.bytes_per_problem;
  const int64_t N = 64;  // n_levels 2, n_cheb 2
  std::atomic<int> completed{0}, failed{0}, wrong{0}, refused{0};
  std::atomic<int64_t> max_submit_us{0};
  double st_srv[11] = {0};
  {
    aifmm::serving::Server<StubOps> srv(budget, slots);
    std::mutex pm;
    std::vector<int64_t> problems;
    std::vector<double> scales;
    auto submitter = [&](int id) {
      std::mt19937_64 rng(seed * 131 + id);
      std::uniform_real_distribution<double> U(0, 1);
      struct Pending {
        int64_t t;
        std::vector<double> b, x;
        double scale;
        int nrhs;
      };
      std::vector<Pending> mine;
      mine.reserve(nreq);
      for (int q = 0; q < nreq; ++q) {
        std::this_thread::sleep_for(std::chrono::microseconds((int)(2000 * U(rng))));
        const auto t0 = std::chrono::steady_clock::now();
        double u = U(rng);
        int64_t pid = -1;
        double scale = 0;
        {
          std::lock_guard<std::mutex> g(pm);
          if (!problems.empty()) {
            size_t k = (size_t)(U(rng) * problems.size()) % problems.size();
            pid = problems[k];
            scale = scales[k];
          }
        }
        if (pid < 0 || u < 0.25) {
          aifmm_ls_params p{};
          p.n_cheb = 2;
          p.degree = 1;
          p.n_levels = 2;
          p.tol_pow = 6;
          p.kappa = 1.0 + id + 0.001 * q;
          int64_t t = 0;
          const int64_t np = srv.submit_factor(p, (int)(U(rng) < 0.8), &t);
          mine.push_back(Pending{t, {}, {}, 0, 0});
          std::lock_guard<std::mutex> g(pm);
          problems.push_back(np);
          scales.push_back(p.kappa);
        } else if (u < 0.30) {
          // a release frees the problem in this thread: not a submission
          try {
            srv.release(pid);
          } catch (...) {
          }
          {
            std::lock_guard<std::mutex> g(pm);
            for (size_t k = 0; k < problems.size(); ++k)
              if (problems[k] == pid) {
                problems.erase(problems.begin() + k);
                scales.erase(scales.begin() + k);
                break;
              }
          }
          continue;
        } else {
          const int nrhs = U(rng) < 0.6 ? 1 : 1 + (int)(U(rng) * 8);
          Pending pd{0, std::vector<double>(2 * N * nrhs), std::vector<double>(2 * N * nrhs, -7.0), scale, nrhs};
          for (size_t k = 0; k < pd.b.size(); ++k) pd.b[k] = 1.0 + (double)((k * 7 + q) % 13);
          pd.b[0] = (double)N;  // tag read by the batch stub
          try {
            pd.t = srv.submit_solve(pid, nrhs, pd.b.data(), pd.x.data(), 1e-9, 10);
            mine.push_back(std::move(pd));
          } catch (...) {
            // released by another submitter meanwhile
          }
        }
        const int64_t us =
            std::chrono::duration_cast<std::chrono::microseconds>(std::chrono::steady_clock::now() - t0).count();
        int64_t m = max_submit_us.load();
        while (us > m && !max_submit_us.compare_exchange_weak(m, us)) {
        }
      }
      for (auto& pd : mine) {
        double info[9];
        std::string err;
        if (srv.wait(pd.t, info, err)) {
          // a solve queued before another submitter released its problem
          if (err.find("released") != std::string::npos) ++refused;
          else ++failed;
          continue;
        }
        ++completed;
        if (pd.nrhs)
          for (size_t k = 0; k < pd.b.size(); ++k)
            if (pd.x[k] != pd.b[k] * pd.scale) {
              ++wrong;
              break;
            }
      }
    };
    std::vector<std::thread> th;
    for (int i = 0; i < nsub; ++i) th.emplace_back(submitter, i);
    for (auto& t : th) t.join();
    srv.stats(st_srv);
  }
  out[0] = completed;
  out[1] = failed;
  out[2] = max_submit_us.load() / 1000.0;
  out[3] = (double)peak_bytes.load();
  out[4] = (double)budget;
  out[5] = st_srv[3];
  out[6] = violations;
  out[7] = st_srv[10];
  out[8] = wrong;
  out[9] = builds;
  out[10] = destroys;
  out[11] = (double)live_bytes.load();
  out[12] = batch_calls;
  out[13] = refused;
  out[14] = peak_building;
  return 0;
}
}
