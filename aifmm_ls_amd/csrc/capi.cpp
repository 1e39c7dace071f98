// C ABI (include/aifmm_ls.h) over the device problem and factorisations.
#include <chrono>
#include <cmath>
#include <cstring>
#include <functional>
#include <memory>
#include <exception>
#include <string>
#include <thread>
#include <vector>
#include <algorithm>
#include <type_traits>

#include "../../include/aifmm_ls.h"
#include "factor.h"
#include "merge.h"
#include "prof.h"
#include "krylov.h"
#include "h2.h"

using namespace aifmm;

struct aifmm_ls_problem {
  std::unique_ptr<Problem> pb;
  std::unique_ptr<Factor<double2>> f64;
  std::unique_ptr<Factor<float2>> f32;
  std::unique_ptr<H2Op> h2;
  std::unique_ptr<Refiner> ref;
  std::unique_ptr<BatchRefiner> bref;  // batches led by this handle
  // merged launch sequences of the last batch led by this handle, reused
  // while its members and their programs stay the same
  struct BatchPlan {
    std::vector<uint64_t> key;  // ids of the members' sweep and matvec programs
    std::unique_ptr<Arena> arena;
    std::unique_ptr<Program<double2>> s64, mat;
    std::unique_ptr<Program<float2>> s32;
    NearDesc* near = nullptr;  // the members' near fields (one launch)
    int near_boxes = 0;
  };
  std::unique_ptr<BatchPlan> bplan;
  int refine_method = 0;
  int precision = -1;
  double l2mf_eps = 1e-12, base_eps = 0;  // zero-pivot epsilons of the next factorisation
  double gate_thr = 1e4;                  // growth gate of the next complex64 factorisation (aifmm_ls_set_growth_gate)
  Arena sc{size_t(16) << 20};
  Arena io{size_t(64) << 20};
  double2* rhs = nullptr;
  double2* sol = nullptr;
  int nrhs = 0;
  double t_solve = 0;
};

static thread_local std::string g_err;

namespace aifmm {
void set_last_error(const std::string& msg) { g_err = msg; }
}  // namespace aifmm

extern "C" {
static int solve_batch_impl(aifmm_ls_problem* const* hs, int n, const double* const* b, double* const* x, double tol,
                            int max_refine, double* relres, double* stats, bool tune);
}

namespace aifmm {
// The problem behind a handle (for entry points defined in other files).
Problem* problem_of(aifmm_ls_problem* h) { return h ? h->pb.get() : nullptr; }

// The batch solve without launch tuning of its merged sequences (serving loop:
// its batches change members from call to call).
int solve_batch_untuned(aifmm_ls_problem* const* hs, int n, const double* const* b, double* const* x, double tol,
                        int max_refine, double* relres, double* stats) {
  return solve_batch_impl(hs, n, b, x, tol, max_refine, relres, stats, false);
}
}  // namespace aifmm

#define GUARD(...)                          \
  try {                                     \
    aifmm::ApiScope api_scope_;             \
    __VA_ARGS__;                            \
    return 0;                               \
  } catch (const std::exception& e) {      \
    g_err = e.what();                       \
    return -1;                              \
  } catch (...) {                           \
    g_err = "unknown error";                \
    return -2;                              \
  }

static double now_s() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

extern "C" {

const char* aifmm_ls_last_error(void) { return g_err.c_str(); }

static aifmm_ls_problem* create_problem(const aifmm_ls_params* p, const std::string& precomp_dir,
                                        int adaptive = 0, int tree_adaptivity = 1) {
  if (!p) throw std::runtime_error("null argument");
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) throw std::runtime_error("no HIP device available");
  Params P;
  P.n_cheb = p->n_cheb;
  P.degree = p->degree;
  P.n_levels = p->n_levels;
  P.tol_pow = p->tol_pow;
  P.qchoice = p->qchoice;
  P.quantize = p->quantize_tables;
  P.kappa = p->kappa;
  P.L = p->L;
  if (p->bumps && p->n_bumps > 0) P.bumps.assign(p->bumps, p->bumps + 4 * p->n_bumps);
  P.adaptive = adaptive;
  P.tree_adaptivity = tree_adaptivity;
  if (P.n_cheb < 1 || P.degree < 1 || P.degree * (P.degree + 1) / 2 > P.n_cheb * P.n_cheb)
    throw std::runtime_error("invalid n_cheb/degree");
  auto* h = new aifmm_ls_problem;
  try {
    h->pb.reset(new Problem(P, precomp_dir));
  } catch (...) {
    delete h;
    throw;
  }
  h->io.set_chunk(Arena::chunk_for(h->pb->N, 256, 4 << 20, 64 << 20));
  h->sc.set_chunk(Arena::chunk_for(h->pb->N, 64, 4 << 20, 16 << 20));
  return h;
}

int aifmm_ls_create(const aifmm_ls_params* p, aifmm_ls_problem** out) {
  GUARD({
    if (!out) throw std::runtime_error("null argument");
    *out = create_problem(p, std::string());
  })
}

int aifmm_ls_create_cached(const aifmm_ls_params* p, const char* dir, aifmm_ls_problem** out) {
  GUARD({
    if (!out || !dir || !*dir) throw std::runtime_error("null argument");
    *out = create_problem(p, dir);
  })
}

int aifmm_ls_create_adaptive(const aifmm_ls_params* p, int tree_adaptivity, aifmm_ls_problem** out) {
  GUARD({
    if (!out) throw std::runtime_error("null argument");
    *out = create_problem(p, std::string(), 1, tree_adaptivity);
  })
}

int aifmm_ls_leaves(aifmm_ls_problem* h, int* levels, int64_t* boxes, double* centres, int* n) {
  GUARD({
    if (!h || !n) throw std::runtime_error("null argument");
    const Problem& P = *h->pb;
    if (!P.P.adaptive) throw std::runtime_error("leaves: the problem has the uniform tree");
    if (levels || boxes || centres) {
      if (*n < P.nleaf) throw std::runtime_error("leaf buffers too small");
      for (int k = 0; k < P.nleaf; ++k) {
        if (levels) levels[k] = P.leaves[k].level;
        if (boxes) boxes[k] = P.leaves[k].num;
        if (centres) centres[2 * k] = P.leaves[k].cx, centres[2 * k + 1] = P.leaves[k].cy;
      }
    }
    *n = P.nleaf;
  })
}

int aifmm_ls_write_precomputations(aifmm_ls_problem* h, const char* dir, double* seconds) {
  GUARD({
    if (!h || !dir || !*dir) throw std::runtime_error("null argument");
    double t = h->pb->write_precomputations(dir);
    if (seconds) *seconds = t;
  })
}

int aifmm_ls_table(aifmm_ls_problem* h, int family, int level, int n, double* out) {
  GUARD({
    if (!h || !out) throw std::runtime_error("null argument");
    if (family < -1 || family > 3) throw std::runtime_error("family must be -1 (far) or 0..3 (neighbours)");
    if (level < 0 || level > aifmm::FAR_LEVELS) throw std::runtime_error("level out of range");
    auto t = h->pb->quad_tables({aifmm::TableSpec{family, level, n}}, nullptr);
    std::memcpy(out, t[0].data(), t[0].size() * 16);
  })
}

void aifmm_ls_destroy(aifmm_ls_problem* h) {
  aifmm::ApiScope api_scope_;
  delete h;
}

int64_t aifmm_ls_size(const aifmm_ls_problem* h) { return h ? h->pb->N : -1; }

int aifmm_ls_points(aifmm_ls_problem* h, double* xy, double* q) {
  GUARD({
    for (int64_t i = 0; i < h->pb->N; ++i) {
      if (xy) xy[2 * i] = h->pb->px[i], xy[2 * i + 1] = h->pb->py[i];
      if (q) q[i] = h->pb->q[i];
    }
  })
}

int aifmm_ls_tables(aifmm_ls_problem* h, double* M, double* T, int* capped) {
  GUARD({
    h->pb->need_uniform("tables of the leaf level");
    if (M) std::memcpy(M, h->pb->M_h.data(), h->pb->M_h.size() * 16);
    if (T) std::memcpy(T, h->pb->T_h.data(), h->pb->T_h.size() * 16);
    if (capped) *capped = h->pb->quad_capped;
  })
}

int aifmm_ls_entries(aifmm_ls_problem* h, int64_t n, const int64_t* I, const int64_t* J, double* out) {
  GUARD({
    for (int64_t t = 0; t < n; ++t)
      if (I[t] < 0 || I[t] >= h->pb->N || J[t] < 0 || J[t] >= h->pb->N) throw std::runtime_error("index out of range");
    h->pb->entries(n, I, J, (double2*)out);
  })
}

int aifmm_ls_rhs_planewave(aifmm_ls_problem* h, double theta, double* b) {
  GUARD({
    const Problem& P = *h->pb;
    double k = P.P.kappa;
    for (int64_t i = 0; i < P.N; ++i) {
      // -kappa^2 q(x) exp(i kappa d.x) (LS_functions.hpp:113-120; d = (1,0) at theta = 0)
      double ph = theta == 0.0 ? k * P.px[i] : k * (cos(theta) * P.px[i] + sin(theta) * P.py[i]);
      double a = -1.0 * k * k * P.q[i];
      b[2 * i] = a * cos(ph);
      b[2 * i + 1] = a * sin(ph);
    }
  })
}

int aifmm_ls_assemble(aifmm_ls_problem* h) { GUARD({ h->pb->assemble(); }) }

int aifmm_ls_set_tolerance(aifmm_ls_problem* h, int tol_pow) {
  GUARD({
    if (h->pb->assembled) throw std::runtime_error("the tolerance is fixed once the operator is assembled");
    if (tol_pow < 1 || tol_pow > 15) throw std::runtime_error("tol_pow must be in [1, 15]");
    h->pb->P.tol_pow = tol_pow;
  })
}

int aifmm_ls_ranks(aifmm_ls_problem* h, int level, int* ranks) {
  GUARD({
    Problem& P = *h->pb;
    if (!P.assembled) throw std::runtime_error("assemble first");
    if (level < 2 || level > P.L) throw std::runtime_error("level out of range");
    for (size_t k = 0; k < P.sk[level].size(); ++k) ranks[k] = P.sk[level][k].r;
  })
}

int aifmm_ls_skeleton(aifmm_ls_problem* h, int level, int box, int which, int64_t* out, int* n) {
  GUARD({
    Problem& P = *h->pb;
    if (!P.assembled) throw std::runtime_error("assemble first");
    if (level < 2 || level > P.L || box < 0 || box >= (int)P.sk[level].size()) throw std::runtime_error("bad box");
    const SkelBox& B = P.sk[level][box];
    const std::vector<int64_t>& v = which == 0 ? B.in_chk : which == 1 ? B.in_chg : which == 2 ? B.out_chk : B.out_chg;
    if ((int)v.size() > *n) throw std::runtime_error("skeleton buffer too small");
    std::memcpy(out, v.data(), v.size() * 8);
    *n = (int)v.size();
  })
}

int aifmm_ls_factor(aifmm_ls_problem* h, int precision) {
  GUARD({
    if (precision != 0 && precision != 1) throw std::runtime_error("precision must be 0 (c128) or 1 (c64)");
    h->pb->assemble();
    // no usable factor until the new one is complete: a failed factorisation
    // leaves the handle unfactored (solves then report "factor first")
    h->precision = -1;
    h->f64.reset();
    h->f32.reset();
    if (precision == 0) {
      std::unique_ptr<Factor<double2>> f(new Factor<double2>(*h->pb));
      f->l2mf_eps = h->l2mf_eps;
      f->base_eps = h->base_eps;
      f->factor();
      h->f64 = std::move(f);
    } else {
      std::unique_ptr<Factor<float2>> f(new Factor<float2>(*h->pb));
      f->l2mf_eps = h->l2mf_eps;
      f->base_eps = h->base_eps;
      f->gate_thr = h->gate_thr;
      f->factor();
      h->f32 = std::move(f);
    }
    h->precision = precision;
  })
}

// One host thread per problem, all on one stream, their batched launches
// merged round by round (MergeGroup): the eliminations of every problem's
// level step, colour by colour, go out as one launch sequence.  Problems of
// different depth start together at their leaf levels.  Each factorisation
// equals the one aifmm_ls_factor makes (every descriptor is computed alone).
int aifmm_ls_factor_batch(aifmm_ls_problem* const* hs, int n, int precision, double* info) {
  GUARD({
    if (precision != 0 && precision != 1) throw std::runtime_error("precision must be 0 (c128) or 1 (c64)");
    if (n < 1 || !hs) throw std::runtime_error("empty batch");
    for (int i = 0; i < n; ++i) {
      if (!hs[i]) throw std::runtime_error("null handle in batch");
      for (int k = 0; k < i; ++k)
        if (hs[k] == hs[i]) throw std::runtime_error("a problem appears twice in the batch");
    }
    const double t0 = now_s();
    std::vector<std::exception_ptr> err(n);
    auto each = [&](const std::function<void(int)>& fn) {
      std::vector<std::thread> th;
      th.reserve(n);
      for (int i = 0; i < n; ++i) th.emplace_back([&, i] {
        try {
          fn(i);
        } catch (...) {
          err[i] = std::current_exception();
        }
      });
      for (auto& t : th) t.join();
    };
    // assembly: each problem on its own stream, concurrently
    each([&](int i) {
      ApiScope api;
      hs[i]->pb->assemble();
      HIP_OK(hipStreamSynchronize(hs[i]->pb->stream));
    });
    for (int i = 0; i < n; ++i)
      if (err[i]) std::rethrow_exception(err[i]);
    const double t1 = now_s();
    for (int i = 0; i < n; ++i) {
      hs[i]->precision = -1;
      hs[i]->f64.reset();
      hs[i]->f32.reset();
    }
    MergeCtx mc;
    mc.s = stream_create(hipStreamNonBlocking);
    const hipStream_t gs = mc.s;
    MergeGroup grp(n, &mc, &MergeCtx::round_end);
    each([&](int i) {
      MergeMember member(&grp);  // first: any exit leaves the group
      ApiScope api;              // counted inside the library: no graph capture meanwhile
      aifmm_ls_problem* h = hs[i];
      Problem& pb = *h->pb;
      const hipStream_t own = pb.stream;
      pb.stream = gs;
      struct Back {
        Problem& pb;
        hipStream_t s;
        ~Back() { pb.stream = s; }
      } back{pb, own};
      if (precision == 0) {
        std::unique_ptr<Factor<double2>> f(new Factor<double2>(pb));
        f->l2mf_eps = h->l2mf_eps;
        f->base_eps = h->base_eps;
        f->factor();
        HIP_OK(hipStreamSynchronize(gs));
        h->f64 = std::move(f);
      } else {
        std::unique_ptr<Factor<float2>> f(new Factor<float2>(pb));
        f->l2mf_eps = h->l2mf_eps;
        f->base_eps = h->base_eps;
        f->gate_thr = h->gate_thr;
        f->factor();
        HIP_OK(hipStreamSynchronize(gs));
        h->f32 = std::move(f);
      }
      h->precision = precision;
    });
    stream_destroy(gs);
    if (info) {
      info[0] = now_s() - t0;  // wall seconds: assembly + factorisation
      info[1] = now_s() - t1;  // wall seconds of the merged factorisations
      info[2] = (double)grp.rounds();
      info[3] = (double)grp.launches();
      info[4] = grp.launch_s();
      info[5] = grp.gather_s();
    }
    for (int i = 0; i < n; ++i)
      if (err[i]) {
        try {
          std::rethrow_exception(err[i]);
        } catch (const std::exception& e) {
          throw std::runtime_error("batch member " + std::to_string(i) + ": " + e.what());
        }
      }
  })
}

int aifmm_ls_set_growth_gate(aifmm_ls_problem* h, double thr) {
  GUARD({
    if (!h) throw std::runtime_error("null handle");
    if (std::isnan(thr)) throw std::runtime_error("growth gate threshold is NaN");
    h->gate_thr = thr < 0 ? -1 : thr;
  })
}

int aifmm_ls_set_pivot_eps(aifmm_ls_problem* h, double l2mf_eps, double base_eps) {
  GUARD({
    if (!h) throw std::runtime_error("null handle");
    if (!(l2mf_eps >= 0 && l2mf_eps < 1) || !(base_eps >= 0 && base_eps < 1))
      throw std::runtime_error("zero-pivot epsilons must lie in [0, 1)");
    h->l2mf_eps = l2mf_eps;
    h->base_eps = base_eps;
  })
}

int aifmm_ls_factor_trace(aifmm_ls_problem* h, int which, int* out, int cap, int* n) {
  GUARD({
    if (!h || !n) throw std::runtime_error("null argument");
    if (h->precision < 0) throw std::runtime_error("factor first");
    const FactorStats& st = h->precision == 0 ? h->f64->st : h->f32->st;
    if (which != 0 && which != 1) throw std::runtime_error("which must be 0 (recompressions) or 1 (zeroed pivots)");
    const std::vector<int>& v = which == 0 ? st.rc_trace : st.zero_piv;
    *n = (int)v.size();
    if (out)
      for (int i = 0; i < (int)v.size() && i < cap; ++i) out[i] = v[i];
  })
}

int aifmm_ls_set_rhs(aifmm_ls_problem* h, int nrhs, const double* b) {
  GUARD({
    if (nrhs < 1) throw std::runtime_error("nrhs must be >= 1");
    int64_t N = h->pb->N;
    h->io.reset();
    h->rhs = h->io.get<double2>(N * nrhs);
    h->sol = h->io.get<double2>(N * nrhs);
    h->nrhs = nrhs;
    h2d(h->rhs, b, N * nrhs * 16, h->pb->stream);
  })
}

int aifmm_ls_solve_resident(aifmm_ls_problem* h, double tol, int max_refine, double* stats) {
  GUARD({
    if (h->precision < 0) throw std::runtime_error("factor first");
    if (!h->rhs) throw std::runtime_error("set_rhs first");
    const int nr = h->nrhs;
    auto M = [&](const double2* b, double2* x) {
      if (h->precision == 0) h->f64->solve(b, x, nr);
      else h->f32->solve(b, x, nr);
    };
    double t0 = now_s();
    RefineResult R;
    if (max_refine > 0) {
      if (!h->h2) h->h2.reset(new H2Op(*h->pb));
      if (!h->ref) h->ref.reset(new Refiner(*h->pb, *h->h2));
      R = h->ref->run(M, h->rhs, h->sol, nr, tol, max_refine, h->refine_method);
    } else {
      M(h->rhs, h->sol);
    }
    HIP_OK(hipStreamSynchronize(h->pb->stream));
    h->t_solve = now_s() - t0;
    if (stats) {
      stats[0] = h->t_solve;
      stats[1] = R.steps;
      stats[2] = R.relres;
    }
  })
}

int aifmm_ls_set_refinement(aifmm_ls_problem* h, int method) {
  GUARD({
    if (method != 0 && method != 1) throw std::runtime_error("refinement method must be 0 (FGMRES) or 1 (Richardson)");
    h->refine_method = method;
  })
}

int aifmm_ls_get_solution(aifmm_ls_problem* h, double* x) {
  GUARD({
    if (!h->sol) throw std::runtime_error("no solution");
    d2h(x, h->sol, h->pb->N * h->nrhs * 16, h->pb->stream);
  })
}

// The batch solve; tune: launch-tune newly merged launch sequences on their
// first run (worth it for batches that repeat, not for one-off ones)
static int solve_batch_impl(aifmm_ls_problem* const* hs, int n, const double* const* b, double* const* x, double tol,
                            int max_refine, double* relres, double* stats, bool tune) {
  GUARD({
    if (n < 1 || !hs || !b || !x) throw std::runtime_error("empty batch");
    const int prec = hs[0]->precision;
    if (prec < 0) throw std::runtime_error("factor first");
    for (int i = 0; i < n; ++i) {
      if (!hs[i] || hs[i]->precision != prec) throw std::runtime_error("batch problems must share the factor precision");
      for (int j = 0; j < i; ++j)
        if (hs[j] == hs[i]) throw std::runtime_error("a problem appears twice in the batch");
    }
    const double t0 = now_s();
    // one stream for the whole batch; the problems' own streams are drained first
    hipStream_t s = hs[0]->pb->stream;
    for (int i = 0; i < n; ++i) HIP_OK(hipStreamSynchronize(hs[i]->pb->stream));
    struct Vec2 {
      double2 *b, *x;
      int64_t N;
    };
    std::vector<Vec2> v(n);
    for (int i = 0; i < n; ++i) {
      aifmm_ls_problem* h = hs[i];
      const int64_t N = h->pb->N;
      h->io.reset();
      v[i] = Vec2{h->io.get<double2>(N), h->io.get<double2>(N), N};
      h->rhs = v[i].b;
      h->sol = v[i].x;
      h->nrhs = 1;
      h2d(v[i].b, b[i], N * 16, s);
    }
    // merged sweep and matvec launch sequences (cached on the leading handle)
    std::vector<const Program<double2>*> ps64, pm;
    std::vector<const Program<float2>*> ps32;
    std::vector<uint64_t> key;
    for (int i = 0; i < n; ++i) {
      aifmm_ls_problem* h = hs[i];
      if (!h->h2) h->h2.reset(new H2Op(*h->pb));
      if (prec == 0) ps64.push_back(&h->f64->sweep_program(1));
      else ps32.push_back(&h->f32->sweep_program(1));
      pm.push_back(&h->h2->program(1));
      key.push_back(prec == 0 ? ps64.back()->id() : ps32.back()->id());
      key.push_back(pm.back()->id());
    }
    auto& bp = hs[0]->bplan;
    if (!bp || bp->key != key) {
      bp.reset(new aifmm_ls_problem::BatchPlan());
      bp->key = key;
      bp->arena.reset(new Arena(size_t(16) << 20));
      bp->mat.reset(new Program<double2>());
      bp->mat->merge_from(pm);
      bp->mat->finalize(*bp->arena, s);
      std::vector<NearDesc> nd(n);
      for (int i = 0; i < n; ++i) {
        nd[i] = hs[i]->h2->near_desc(1);
        bp->near_boxes = std::max(bp->near_boxes, nd[i].nbox);
      }
      bp->near = upload(*bp->arena, nd, s);
      if (prec == 0) {
        bp->s64.reset(new Program<double2>());
        bp->s64->merge_from(ps64);
        bp->s64->finalize(*bp->arena, s);
      } else {
        bp->s32.reset(new Program<float2>());
        bp->s32->merge_from(ps32);
        bp->s32->finalize(*bp->arena, s);
      }
    }
    Program<double2>& mat = *bp->mat;
    // lockstep flexible GMRES; preconditioner and operator of all members in
    // merged launch sequences
    // (the merged programs are launch-tuned on their first run: the loads
    // are issued again before every timed run)
    auto Mall = [&](const std::vector<const double2*>& src, const std::vector<double2*>& dst) {
      auto load = [&] {
        for (int i = 0; i < n; ++i) {
          if (prec == 0) hs[i]->f64->sweep_load(src[i], 1, s);
          else hs[i]->f32->sweep_load(src[i], 1, s);
        }
      };
      load();
      const std::function<void()> reload = tune ? std::function<void()>(load) : std::function<void()>();
      if (prec == 0) bp->s64->run(1, s, reload);
      else bp->s32->run(1, s, reload);
      for (int i = 0; i < n; ++i) {
        if (prec == 0) hs[i]->f64->sweep_store(dst[i], 1, s);
        else hs[i]->f32->sweep_store(dst[i], 1, s);
      }
    };
    auto Aall = [&](const std::vector<const double2*>& src, const std::vector<double2*>& dst) {
      auto load = [&] {
        for (int i = 0; i < n; ++i) hs[i]->h2->load(src[i], 1, s);
      };
      load();
      mat.run(1, s, tune ? std::function<void()>(load) : std::function<void()>());
      launch_near(n, bp->near_boxes, bp->near, s);
      for (int i = 0; i < n; ++i) hs[i]->h2->store(dst[i], 1, s);
    };
    std::vector<BatchVec> bv(n);
    for (int i = 0; i < n; ++i) bv[i] = BatchVec{v[i].N, v[i].b, v[i].x};
    if (!hs[0]->bref) hs[0]->bref.reset(new BatchRefiner());
    std::vector<RefineResult> R = hs[0]->bref->run(bv, Mall, Aall, tol, max_refine, s);
    int steps = 0;
    std::vector<double> rel(n, -1.0);
    for (int i = 0; i < n; ++i) {
      rel[i] = R[i].relres;
      steps = std::max(steps, R[i].steps);
    }
    for (int i = 0; i < n; ++i) d2h(x[i], v[i].x, v[i].N * 16, s);
    HIP_OK(hipStreamSynchronize(s));
    const double t = now_s() - t0;
    double worst = 0;
    for (int i = 0; i < n; ++i) {
      if (relres) relres[i] = rel[i];
      worst = std::max(worst, rel[i]);
      hs[i]->t_solve = t;
    }
    if (stats) {
      stats[0] = t;
      stats[1] = steps;
      stats[2] = worst;
    }
  })
}


int aifmm_ls_solve_batch(aifmm_ls_problem* const* hs, int n, const double* const* b, double* const* x, double tol,
                         int max_refine, double* relres, double* stats) {
  return solve_batch_impl(hs, n, b, x, tol, max_refine, relres, stats, true);
}

int aifmm_ls_solve(aifmm_ls_problem* h, int nrhs, const double* b, double* x, double tol, int max_refine,
                   double* stats) {
  int rc = aifmm_ls_set_rhs(h, nrhs, b);
  if (rc) return rc;
  rc = aifmm_ls_solve_resident(h, tol, max_refine, stats);
  if (rc) return rc;
  return aifmm_ls_get_solution(h, x);
}

int aifmm_ls_matvec(aifmm_ls_problem* h, int nrhs, const double* x, double* y) {
  GUARD({
    h->pb->assemble();
    if (!h->h2) h->h2.reset(new H2Op(*h->pb));
    int64_t N = h->pb->N;
    Arena a{size_t(64) << 20};
    double2* dx = a.get<double2>(N * nrhs);
    double2* dy = a.get<double2>(N * nrhs);
    h2d(dx, x, N * nrhs * 16, h->pb->stream);
    h->h2->apply(dx, dy, nrhs);
    HIP_OK(hipStreamSynchronize(h->pb->stream));
    d2h(y, dy, N * nrhs * 16, h->pb->stream);
  })
}

int aifmm_ls_dense_matvec(aifmm_ls_problem* h, int nrhs, const double* x, double* y) {
  GUARD({
    if (!h || !x || !y) throw std::runtime_error("null argument");
    if (nrhs < 1) throw std::runtime_error("nrhs must be >= 1");
    Problem& P = *h->pb;
    const int64_t N = P.N;
    const int nch = dmv_chunks(N, P.nleaf);
    const size_t ns = dmv_scratch(N, P.nleaf, nch);
    Arena a{(2 * (size_t)N + ns) * sizeof(double2) + (size_t(1) << 20)};
    double2* dx = a.get<double2>(N);
    double2* dy = a.get<double2>(N);
    double2* sc = a.get<double2>(ns);
    for (int c = 0; c < nrhs; ++c) {
      h2d(dx, x + 2 * (size_t)N * c, (size_t)N * 16, P.stream);
      launch_dense_matvec(P.ctx_phi(), N, dx, dy, sc, nch, P.stream);
      d2h(y + 2 * (size_t)N * c, dy, (size_t)N * 16, P.stream);
    }
  })
}

int aifmm_ls_set_find_phi(aifmm_ls_problem* h, int on) {
  GUARD({
    if (!h) throw std::runtime_error("null handle");
    h->pb->find_phi = on ? 1 : 0;
  })
}

int aifmm_ls_incident_field(aifmm_ls_problem* h, double theta, double* u) {
  GUARD({
    if (!h || !u) throw std::runtime_error("null argument");
    const Problem& P = *h->pb;
    const double k = P.P.kappa;
    for (int64_t i = 0; i < P.N; ++i) {
      // exp(i kappa d.x) (LS_functions.hpp:113-116; d = (1, 0) at theta = 0)
      const double ph = theta == 0.0 ? k * P.px[i] : k * (cos(theta) * P.px[i] + sin(theta) * P.py[i]);
      u[2 * i] = cos(ph);
      u[2 * i + 1] = sin(ph);
    }
  })
}

int aifmm_ls_total_field(aifmm_ls_problem* h, double theta, const double* phi, double* u) {
  GUARD({
    if (!h || !phi || !u) throw std::runtime_error("null argument");
    Problem& P = *h->pb;
    const int64_t N = P.N;
    const int keep = P.find_phi;
    P.find_phi = 0;
    std::vector<double> us(2 * (size_t)N);
    const int rc = aifmm_ls_dense_matvec(h, 1, phi, us.data());
    P.find_phi = keep;
    if (rc) throw std::runtime_error(aifmm_ls_last_error());
    if (aifmm_ls_incident_field(h, theta, u)) throw std::runtime_error(aifmm_ls_last_error());
    for (size_t t = 0; t < us.size(); ++t) u[t] += us[t];
  })
}

int aifmm_ls_device_bytes(aifmm_ls_problem* h, int64_t* bytes) {
  GUARD({
    if (!h || !bytes) throw std::runtime_error("null argument");
    size_t b = h->pb->device_bytes() + h->sc.reserved() + h->io.reserved();
    if (h->f64) b += h->f64->device_bytes();
    if (h->f32) b += h->f32->device_bytes();
    if (h->h2) b += h->h2->device_bytes();
    if (h->ref) b += h->ref->device_bytes();
    *bytes = (int64_t)b;
  })
}

int aifmm_ls_stats(aifmm_ls_problem* h, double* out, int n) {
  GUARD({
    double v[21] = {0};
    Problem& P = *h->pb;
    v[0] = P.t_tables;
    v[1] = P.t_assemble;
    const FactorStats* st = h->f64 ? &h->f64->st : h->f32 ? &h->f32->st : nullptr;
    if (st) {
      v[2] = st->t_factor;
      v[3] = st->flops;
      v[4] = (double)st->bytes;
      v[5] = st->base_n;
      v[6] = st->max_rank;
      v[8] = st->recompressions;
      v[12] = st->ill_boxes;
      v[15] = st->dense_level;
      v[16] = st->dense_n;
      v[17] = st->base_rdiag_ratio;
      v[18] = (double)(st->zero_piv.size() / 4);
      v[19] = st->piv_c64;
      v[20] = st->piv_fp64;
    }
    if (h->f64) v[7] = (double)h->f64->solve_bytes(1);
    if (h->f32) v[7] = (double)h->f32->solve_bytes(1);
    v[9] = (double)P.operator_bytes();
    v[10] = h->t_solve;
    v[11] = P.t_nca;
    // the latest launch tuning of the active factorisation's sweep
    auto tuned = [&](const auto* tr) {
      if (!tr || tr->default_ms <= 0) return;
      v[13] = (tr->adopted ? tr->tuned_ms : tr->default_ms) / tr->default_ms;
      v[14] = tr->adopted ? tr->changed : 0;
    };
    if (h->precision == 0 && h->f64) tuned(h->f64->sweep_tune());
    if (h->precision == 1 && h->f32) tuned(h->f32->sweep_tune());
    for (int i = 0; i < n && i < 21; ++i) out[i] = v[i];
  })
}

}  // extern "C"

// Kernel self-test hook: runs one batched dense kernel on host inputs.
// op 0: C = A(m x k) B(k x n)          op 1: C = A(m x m)^{-1} B(m x n) (LU)
// op 2: CPQR of A (m x n) at threshold thr -> iout[0] = rank, C = Q[:, :k]
// op 3: C = A(m x m)^{-1} B(m x n) by CPQR + ColPivHouseholderQR::solve,
//       iout[0] = number of nonzero pivots
// op 4 (complex64 only, op 14): the pivot-block growth gate (below)
// op 5: CPQR of A (m x n) at threshold thr by the kernel k & 1 (0 k_cpqr,
//       1 k_cpqr_reg), with early stop if k & 2: C = the overwritten A then
//       tau, iout = rank, nonzero pivots, perm[0..n)
// op +10: the same in complex64.  Column-major interleaved complex.
template <class T>
static void selftest_t(int op, int m, int n, int k, const double* A, const double* B, double* C, int* iout,
                       double thr) {
  Arena a{size_t(64) << 20}, sc{size_t(16) << 20};
  hipStream_t s = stream_create(hipStreamDefault);
  auto up = [&](const double* h, int r, int c) {
    Blk<double2> d = new_blk<double2>(a, r, c);
    HIP_OK(hipMemcpy(d.p, h, (size_t)r * c * 16, hipMemcpyHostToDevice));
    Blk<T> t = new_blk<T>(a, r, c);
    std::vector<CvtDesc<T>> cv{CvtDesc<T>{d.p, t.p, (int64_t)r * c}};
    auto* dd = upload(sc, cv, s);
    launch_cvt<T>(1, dd, s, max_len(cv));
    HIP_OK(hipStreamSynchronize(s));
    return t;
  };
  auto down = [&](const Blk<T>& t, double* h) {
    Blk<double2> d = new_blk<double2>(a, t.r, t.c);
    if (std::is_same<T, double2>::value) HIP_OK(hipMemcpy(d.p, t.p, t.size() * 16, hipMemcpyDeviceToDevice));
    else launch_f2d((const float2*)t.p, d.p, t.size(), s);
    HIP_OK(hipStreamSynchronize(s));
    HIP_OK(hipMemcpy(h, d.p, t.size() * 16, hipMemcpyDeviceToHost));
  };
  if (op == 0) {
    Blk<T> dA = up(A, m, k), dB = up(B, k, n), dC = new_blk<T>(a, m, n);
    GemmBatch<T> g;
    g.add(dA, 0, dB, 0, dC, make_double2(1, 0), make_double2(0, 0));
    g.run(sc, s);
    down(dC, C);
  } else if (op == 1) {
    Blk<T> dA = up(A, m, m), dB = up(B, m, n);
    int* piv = a.get<int>(m);
    LUBatch<T> lu;
    lu.add(dA, piv);
    lu.run(sc, s);
    SolveBatch<T> sv;
    sv.add(dA, piv, dB);
    sv.run(sc, s);
    down(dB, C);
  } else if (op == 2) {
    using Rl = typename CplxOf<T>::real;
    Blk<T> dA = up(A, m, n);
    int mn = std::min(m, n);
    QRDesc<T> d{};
    d.A = dA.p;
    d.m = m;
    d.n = n;
    d.lda = m;
    d.tau = a.get<T>(mn);
    d.perm = a.get<int>(n);
    d.work = a.get<Rl>(2 * (size_t)n);
    d.thr = thr;
    d.rank_out = a.get<int>(1);
    d.maxpiv_out = nullptr;
    std::vector<QRDesc<T>> v{d};
    launch_cpqr<T>(1, upload(sc, v, s), s, cpqr_shape(m, n));
    Blk<T> Q = new_blk<T>(a, m, k);
    std::vector<QFormDesc<T>> f{QFormDesc<T>{dA.p, d.tau, Q.p, m, n, k, m, m}};
    launch_qform<T>(1, upload(sc, f, s), s);
    HIP_OK(hipStreamSynchronize(s));
    HIP_OK(hipMemcpy(iout, d.rank_out, 4, hipMemcpyDeviceToHost));
    down(Q, C);
  } else if (op == 5) {
    using Rl = typename CplxOf<T>::real;
    const int shape = k & 1;
    if (shape == CPQR_REG && cpqr_shape(m, n) != CPQR_REG)
      throw std::runtime_error("k_cpqr_reg takes at most 36 rows and 256 columns");
    Blk<T> dA = up(A, m, n);
    const int mn = std::min(m, n);
    Blk<T> tau = new_blk<T>(a, std::max(mn, 1), 1);
    QRDesc<T> d{};
    d.A = dA.p;
    d.m = m;
    d.n = n;
    d.lda = m;
    d.tau = tau.p;
    d.perm = a.get<int>(n);
    d.work = a.get<Rl>(2 * (size_t)n);
    d.thr = thr;
    d.early = (k & 2) ? 1 : 0;
    int* outs = a.get<int>(2);
    d.rank_out = outs;
    d.nz_out = outs + 1;
    d.maxpiv_out = nullptr;
    HIP_OK(hipMemsetAsync(tau.p, 0, sizeof(T) * std::max(mn, 1), s));
    std::vector<QRDesc<T>> v{d};
    launch_cpqr<T>(1, upload(sc, v, s), s, shape);
    HIP_OK(hipStreamSynchronize(s));
    down(dA, C);
    if (mn) down(tau.sub(0, 0, mn, 1), C + 2 * (size_t)m * n);
    HIP_OK(hipMemcpy(iout, outs, 8, hipMemcpyDeviceToHost));
    HIP_OK(hipMemcpy(iout + 2, d.perm, 4 * (size_t)n, hipMemcpyDeviceToHost));
  } else if (op == 3) {
    Blk<T> dA = up(A, m, m), dB = up(B, m, n);
    QRFact<T> f = new_qr(a, dA);
    QRBatch<T> qb;
    qb.add(f, sc);
    qb.run(sc, s);
    QRSolveBatch<T> qs;
    qs.add(f, dB, sc);
    qs.run(sc, s);
    HIP_OK(hipStreamSynchronize(s));
    HIP_OK(hipMemcpy(iout, f.nz, 4, hipMemcpyDeviceToHost));
    down(dB, C);
  } else if (op == 4) {
    // complex64 pivot block through the growth gate at threshold thr:
    // C[:m*n] = A^{-1} B, C[m*n].re = the block's estimate, iout[0] = 1 when
    // the block went to FP64
    if constexpr (std::is_same<T, float2>::value) {
      Blk<float2> dA = up(A, m, m), dB = up(B, m, n);
      DenseOps<float2> dn;
      DFact<float2> f;
      dn.factor(a, sc, f, dA, 0, thr);
      dn.run_factor(sc, s);
      dn.solve(sc, f, dB);
      dn.run_solve(sc, s);
      HIP_OK(hipStreamSynchronize(s));
      float est = INFINITY;
      iout[0] = 1;
      if (f.flag) {
        HIP_OK(hipMemcpy(iout, f.flag, 4, hipMemcpyDeviceToHost));
        HIP_OK(hipMemcpy(&est, f.est, 4, hipMemcpyDeviceToHost));
        iout[0] = iout[0] != 0;
      }
      down(dB, C);
      C[2 * (size_t)m * n] = est;
      C[2 * (size_t)m * n + 1] = 0;
    } else {
      throw std::runtime_error("self-test op 4 is complex64 only (op 14)");
    }
  }
  HIP_OK(hipStreamSynchronize(s));
  stream_destroy(s);
}

// One ACA (NCA pivot search) on explicit point lists, as getNodes_* run it.
extern "C" int aifmm_ls_aca(aifmm_ls_problem* h, int tol_pow, int n1, const int64_t* rows, int n2,
                            const int64_t* cols, int* rb, int* cb, double* Ac, double* Ar, int* rank) {
  GUARD({
    Problem& P = *h->pb;
    Arena a{size_t(256) << 20};
    int np = std::min(n1, n2), ns = std::max(n1, n2);
    int64_t* dr = a.get<int64_t>(n1 > 0 ? n1 : 1);
    int64_t* dc = a.get<int64_t>(n2 > 0 ? n2 : 1);
    HIP_OK(hipMemcpy(dr, rows, 8 * n1, hipMemcpyHostToDevice));
    HIP_OK(hipMemcpy(dc, cols, 8 * n2, hipMemcpyHostToDevice));
    AcaProblem p{};
    p.rows = dr;
    p.cols = dc;
    p.np = np;
    p.ns = ns;
    p.prim_rows = n1 < n2;
    p.rmax = np;
    size_t rm = np > 0 ? np : 1;
    p.U = a.get<double2>(rm * rm);
    p.V = a.get<double2>(rm * (ns ? ns : 1));
    p.Craw = a.get<double2>(rm * rm);
    p.Lraw = a.get<double2>(rm * (ns ? ns : 1));
    p.line = a.get<double2>(ns ? ns : 1);
    p.cross = a.get<double2>(rm);
    p.remP = a.get<char>(rm);
    p.remS = a.get<char>(ns ? ns : 1);
    p.pb = a.get<int>(rm);
    p.sb = a.get<int>(rm);
    p.rank_out = a.get<int>(1);
    AcaProblem* dp = a.get<AcaProblem>(1);
    HIP_OK(hipMemcpy(dp, &p, sizeof(p), hipMemcpyHostToDevice));
    launch_aca(P.ctx, 1, dp, pow(10, -1.0 * tol_pow), 1e-12, P.stream);
    HIP_OK(hipStreamSynchronize(P.stream));
    int r = 0;
    HIP_OK(hipMemcpy(&r, p.rank_out, 4, hipMemcpyDeviceToHost));
    *rank = r;
    std::vector<int> pbv(r), sbv(r);
    if (r) {
      HIP_OK(hipMemcpy(pbv.data(), p.pb, 4 * r, hipMemcpyDeviceToHost));
      HIP_OK(hipMemcpy(sbv.data(), p.sb, 4 * r, hipMemcpyDeviceToHost));
    }
    for (int q = 0; q < r; ++q) {
      rb[q] = p.prim_rows ? pbv[q] : sbv[q];
      cb[q] = p.prim_rows ? sbv[q] : pbv[q];
    }
    // Ac (N1 x r): crosses if primary = rows else lines; Ar^T (N2 x r): the other
    std::vector<double2> cr((size_t)r * np), ln((size_t)r * ns);
    if (r) {
      HIP_OK(hipMemcpy(cr.data(), p.Craw, cr.size() * 16, hipMemcpyDeviceToHost));
      HIP_OK(hipMemcpy(ln.data(), p.Lraw, ln.size() * 16, hipMemcpyDeviceToHost));
    }
    const std::vector<double2>& acs = p.prim_rows ? cr : ln;
    const std::vector<double2>& ars = p.prim_rows ? ln : cr;
    if (Ac) std::memcpy(Ac, acs.data(), (size_t)r * n1 * 16);
    if (Ar)
      for (int q = 0; q < r; ++q)
        for (int c = 0; c < n2; ++c) {
          Ar[2 * (q + (size_t)r * c)] = ars[(size_t)q * n2 + c].x;
          Ar[2 * (q + (size_t)r * c) + 1] = ars[(size_t)q * n2 + c].y;
        }
  })
}

// Box-program self-test: nprog programs of m rows in one stage, each
//   y_p = S_p + sum_t alpha_t A_{p,t} X_{p,t}   (phase 1, saved to Y)
//   out_p = B_p y_p                             (phase 2, y read from LDS)
// with alpha_t = 1 for even t and -0.5 + 0.25 i for odd t, run in launch
// shape `shape` (0 = default).  Host arrays column-major interleaved complex:
// A nprog*nterms blocks m x cols, X nprog*nterms blocks cols x nrhs, S and
// Y and Out nprog blocks m x nrhs, Bm nprog blocks m x m.
template <class T>
static void selftest_prog_t(int shape, int nprog, int m, int nterms, int cols, int nrhs, const double* A,
                            const double* X, const double* S, const double* Bm, double* Y, double* Out) {
  Arena a{size_t(64) << 20}, sc{size_t(16) << 20};
  hipStream_t s = stream_create(hipStreamDefault);
  try {
    auto up = [&](const double* h, size_t n) {
      double2* d = a.get<double2>(n);
      HIP_OK(hipMemcpy(d, h, n * 16, hipMemcpyHostToDevice));
      T* t = a.get<T>(n);
      std::vector<CvtDesc<T>> cv{CvtDesc<T>{d, t, (int64_t)n}};
      launch_cvt<T>(1, upload(sc, cv, s), s, max_len(cv));
      return t;
    };
    // vectors stay complex128 (the programs' vector type)
    auto upv = [&](const double* h, size_t n) {
      double2* d = a.get<double2>(n);
      HIP_OK(hipMemcpy(d, h, n * 16, hipMemcpyHostToDevice));
      return d;
    };
    auto down = [&](const double2* d, size_t n, double* h) {
      HIP_OK(hipStreamSynchronize(s));
      HIP_OK(hipMemcpy(h, d, n * 16, hipMemcpyDeviceToHost));
    };
    const size_t mc = (size_t)m * cols, cr = (size_t)cols * nrhs, mr = (size_t)m * nrhs;
    T* dA = up(A, (size_t)nprog * nterms * mc);
    double2* dX = upv(X, (size_t)nprog * nterms * cr);
    double2* dS = upv(S, (size_t)nprog * mr);
    T* dB = up(Bm, (size_t)nprog * m * m);
    double2* dY = a.get<double2>((size_t)nprog * mr);
    double2* dO = a.get<double2>((size_t)nprog * mr);
    Program<T> P(nrhs);  // tiled as a program built for nrhs columns
    P.stage();
    for (int p = 0; p < nprog; ++p) {
      BoxOp<T> op;
      op.m = m;
      op.src = dS + p * mr;
      op.ldsrc = m;
      op.save = dY + p * mr;
      op.ldsave = m;
      for (int t = 0; t < nterms; ++t)
        op.term(dA + ((size_t)p * nterms + t) * mc, m, cols, dX + ((size_t)p * nterms + t) * cr, cols,
                t % 2 ? make_double2(-0.5, 0.25) : make_double2(1, 0));
      auto& o = op.out(dO + p * mr, m, m);
      out_term<T>(o, dB + (size_t)p * m * m, m, m, nullptr, 0, make_double2(1, 0));
      P.add(op);
    }
    P.finalize(a, s);
    P.force_shape(nrhs, shape);
    P.run(nrhs, s);
    HIP_OK(hipStreamSynchronize(s));
    down(dY, (size_t)nprog * mr, Y);
    down(dO, (size_t)nprog * mr, Out);
  } catch (...) {
    stream_destroy(s);
    throw;
  }
  stream_destroy(s);
}

extern "C" int aifmm_ls_selftest_prog(int precision, int shape, int nprog, int m, int nterms, int cols, int nrhs,
                                      const double* A, const double* X, const double* S, const double* Bm, double* Y,
                                      double* Out) {
  GUARD({
    if (nprog < 1 || m < 1 || nterms < 1 || cols < 1 || nrhs < 1) throw std::runtime_error("empty self-test program");
    if (!A || !X || !S || !Bm || !Y || !Out) throw std::runtime_error("null argument");
    if (precision == 0) selftest_prog_t<double2>(shape, nprog, m, nterms, cols, nrhs, A, X, S, Bm, Y, Out);
    else if (precision == 1) selftest_prog_t<float2>(shape, nprog, m, nterms, cols, nrhs, A, X, S, Bm, Y, Out);
    else throw std::runtime_error("precision must be 0 or 1");
  })
}

extern "C" int aifmm_ls_selftest(int op, int m, int n, int k, const double* A, const double* B, double* C, int* iout,
                      double thr) {
  GUARD({
    if (op >= 10) selftest_t<float2>(op - 10, m, n, k, A, B, C, iout, thr);
    else selftest_t<double2>(op, m, n, k, A, B, C, iout, thr);
  })
}



extern "C" int aifmm_ls_selftest_stream(int64_t bytes, int reps, double* ms) {
  GUARD({
    if (bytes < 16 || reps < 1) throw std::runtime_error("stream self-test: bytes >= 16 and reps >= 1");
    const int64_t n = bytes / 16;
    Arena a{size_t(2 * n) * 16 + 4096};
    double2* src = a.get<double2>(n);
    double2* dst = a.get<double2>(n);
    hipStream_t s = stream_create(hipStreamNonBlocking);
    hipEvent_t e0, e1;
    HIP_OK(hipEventCreate(&e0));
    HIP_OK(hipEventCreate(&e1));
    HIP_OK(hipMemsetAsync(src, 0, (size_t)n * 16, s));
    double tot = 0;
    for (int r = 0; r < reps; ++r) {
      HIP_OK(hipEventRecord(e0, s));
      launch_stream_copy(r & 1 ? dst : src, r & 1 ? src : dst, n, s);
      HIP_OK(hipEventRecord(e1, s));
      HIP_OK(hipEventSynchronize(e1));
      float t = 0;
      HIP_OK(hipEventElapsedTime(&t, e0, e1));
      tot += t;
    }
    if (ms) *ms = tot / reps;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    stream_destroy(s);
  })
}

int aifmm_ls_profile_enable(int on) {
  GUARD({
    if (on < 0 || on > 2) throw std::runtime_error("profile mode must be 0, 1 or 2");
    prof_enable(on != 0, on == 2);
  })
}

extern "C" int aifmm_ls_set_tuning(int mode) {
  GUARD({ prog_set_tuning(mode); })
}

extern "C" int aifmm_ls_profile_launches(int reset, int max, double* out, char* names, int names_len) {
  try {
    if (max < 0 || (max > 0 && !out)) throw std::runtime_error("null argument");
    std::vector<LaunchRecord> ls = prof_launches(reset != 0);
    std::string nm;
    int n = 0;
    for (auto& l : ls) {
      if (n >= max) break;
      out[2 * n] = l.bytes;
      out[2 * n + 1] = l.ms;
      nm += std::string(l.name) + "\n";
      ++n;
    }
    if (names && names_len > 0) {
      size_t c = std::min(nm.size(), (size_t)names_len - 1);
      std::memcpy(names, nm.data(), c);
      names[c] = 0;
    }
    return n;
  } catch (const std::exception& e) {
    g_err = e.what();
    return -1;
  }
}

extern "C" int aifmm_ls_profile_read(int reset, int max, double* out, char* names, int names_len) {
  try {
    std::vector<KernelProfile> ks = prof_collect(reset != 0);
    std::string nm;
    int n = 0;
    for (auto& k : ks) {
      if (n >= max) break;
      out[5 * n] = (double)k.launches;
      out[5 * n + 1] = k.ms;
      out[5 * n + 2] = k.bytes;
      out[5 * n + 3] = k.flops;
      out[5 * n + 4] = 0;
      nm += k.name + "\n";
      ++n;
    }
    if (names && names_len > 0) {
      size_t c = std::min(nm.size(), (size_t)names_len - 1);
      std::memcpy(names, nm.data(), c);
      names[c] = 0;
    }
    return n;
  } catch (const std::exception& e) {
    g_err = e.what();
    return -1;
  }
}
