// Serving loop of the C ABI (aifmm_ls_server_*): the scheduling core of
// serve_core.h over real problem handles.  No reference counterpart: the
// reference factorises and solves one problem per process
// (testFMM2D_find_error.cpp:41-176).
#include <atomic>
#include <stdexcept>
#include <string>

#include "../../include/aifmm_ls.h"
#include "prog.h"
#include "serve_core.h"

namespace aifmm {
void set_last_error(const std::string& msg);  // capi.cpp
int solve_batch_untuned(aifmm_ls_problem* const* hs, int n, const double* const* b, double* const* x, double tol,
                        int max_refine, double* relres, double* stats);  // capi.cpp
}

namespace {

struct RealOps {
  using Handle = aifmm_ls_problem*;
  static int build(const aifmm_ls_params& p, int precision, Handle& h, std::string& err) {
    h = nullptr;
    if (aifmm_ls_create(&p, &h) || aifmm_ls_assemble(h) || aifmm_ls_factor(h, precision)) {
      err = aifmm_ls_last_error();
      if (h) aifmm_ls_destroy(h);
      h = nullptr;
      return -1;
    }
    return 0;
  }
  static int solve(Handle h, int nrhs, const double* b, double* x, double tol, int max_refine, double* st,
                   std::string& err) {
    if (aifmm_ls_solve(h, nrhs, b, x, tol, max_refine, st)) {
      err = aifmm_ls_last_error();
      return -1;
    }
    return 0;
  }
  static int solve_batch(Handle* hs, int n, const double* const* b, double* const* x, double tol, int max_refine,
                         double* relres, double* st, std::string& err) {
    if (aifmm::solve_batch_untuned(hs, n, b, x, tol, max_refine, relres, st)) {
      err = aifmm_ls_last_error();
      return -1;
    }
    return 0;
  }
  static int64_t bytes(Handle h) {
    int64_t b = 0;
    return aifmm_ls_device_bytes(h, &b) ? 0 : b;
  }
  static void destroy(Handle h) { aifmm_ls_destroy(h); }
};

using ServerImpl = aifmm::serving::Server<RealOps>;
std::atomic<int> live_servers{0};

}  // namespace

struct aifmm_ls_server {
  ServerImpl impl;
  aifmm_ls_server(int64_t budget, int slots) : impl(budget, slots) {}
};

extern "C" {

#define SGUARD(...)                         \
  try {                                     \
    __VA_ARGS__;                            \
    return 0;                               \
  } catch (const std::exception& ex) {     \
    aifmm::set_last_error(ex.what());       \
    return -1;                              \
  } catch (...) {                           \
    aifmm::set_last_error("unknown error"); \
    return -1;                              \
  }

int aifmm_ls_server_create(int64_t budget_bytes, int slots, aifmm_ls_server** out) {
  SGUARD({
    if (!out) throw std::runtime_error("null argument");
    if (budget_bytes <= 0) throw std::runtime_error("device memory budget must be positive");
    if (slots < 1 || slots > 16) throw std::runtime_error("slots must be in [1, 16]");
    // sweeps launch directly while a server exists: its slots free memory and
    // synchronise streams while others would capture
    if (live_servers++ == 0) aifmm::prog_graphs_hold(true);
    try {
      *out = new aifmm_ls_server(budget_bytes, slots);
    } catch (...) {
      if (--live_servers == 0) aifmm::prog_graphs_hold(false);
      throw;
    }
  })
}

int aifmm_ls_server_factor(aifmm_ls_server* s, const aifmm_ls_params* p, int precision, int64_t* problem,
                           int64_t* ticket) {
  SGUARD({
    if (!s || !p || !problem) throw std::runtime_error("null argument");
    if (precision != 0 && precision != 1) throw std::runtime_error("precision must be 0 (complex128) or 1 (complex64)");
    if (p->n_levels < 2 || p->n_levels > 9) throw std::runtime_error("n_levels must be in [2, 9]");
    *problem = s->impl.submit_factor(*p, precision, ticket);
  })
}

int aifmm_ls_server_solve(aifmm_ls_server* s, int64_t problem, int nrhs, const double* b, double* x, double tol,
                          int max_refine, int64_t* ticket) {
  SGUARD({
    if (!s || !b || !x || !ticket) throw std::runtime_error("null argument");
    *ticket = s->impl.submit_solve(problem, nrhs, b, x, tol, max_refine);
  })
}

int aifmm_ls_server_wait(aifmm_ls_server* s, int64_t ticket, double* info) {
  if (!s) {
    aifmm::set_last_error("null server");
    return -1;
  }
  double v[9] = {0};
  std::string err;
  const int rc = s->impl.wait(ticket, v, err);
  if (info)
    for (int i = 0; i < 9; ++i) info[i] = v[i];
  if (rc) aifmm::set_last_error(err);
  return rc;
}

int aifmm_ls_server_release(aifmm_ls_server* s, int64_t problem) {
  SGUARD({
    if (!s) throw std::runtime_error("null server");
    s->impl.release(problem);
  })
}

int aifmm_ls_server_stats(aifmm_ls_server* s, double* out, int n) {
  SGUARD({
    if (!s) throw std::runtime_error("null server");
    double v[11] = {0};
    s->impl.stats(v);
    for (int i = 0; i < n && i < 11; ++i) out[i] = v[i];
  })
}

int aifmm_ls_server_destroy(aifmm_ls_server* s) {
  if (!s) return 0;
  delete s;
  if (--live_servers == 0) aifmm::prog_graphs_hold(false);
  return 0;
}

}  // extern "C"
