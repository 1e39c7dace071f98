// C ABI (include/aifmm_ls.h) over the device problem and factorisations.
#include <chrono>
#include <cmath>
#include <cstring>
#include <memory>
#include <string>

#include "../../include/aifmm_ls.h"
#include "factor.h"

using namespace aifmm;

struct aifmm_ls_problem {
  std::unique_ptr<Problem> pb;
  std::unique_ptr<Factor<double2>> f64;
  std::unique_ptr<Factor<float2>> f32;
  int precision = -1;
  Arena io{size_t(64) << 20};
  double2* rhs = nullptr;
  double2* sol = nullptr;
  int nrhs = 0;
  double t_solve = 0;
};

static thread_local std::string g_err;

#define GUARD(...)                          \
  try {                                     \
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

int aifmm_ls_create(const aifmm_ls_params* p, aifmm_ls_problem** out) {
  GUARD({
    if (!p || !out) throw std::runtime_error("null argument");
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
    if (P.n_cheb < 1 || P.degree < 1 || P.degree * (P.degree + 1) / 2 > P.n_cheb * P.n_cheb)
      throw std::runtime_error("invalid n_cheb/degree");
    auto* h = new aifmm_ls_problem;
    try {
      h->pb.reset(new Problem(P));
    } catch (...) {
      delete h;
      throw;
    }
    *out = h;
  })
}

void aifmm_ls_destroy(aifmm_ls_problem* h) { delete h; }

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
    h->f64.reset();
    h->f32.reset();
    if (precision == 0) {
      h->f64.reset(new Factor<double2>(*h->pb));
      h->f64->factor();
    } else {
      h->f32.reset(new Factor<float2>(*h->pb));
      h->f32->factor();
    }
    h->precision = precision;
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
    HIP_OK(hipMemcpy(h->rhs, b, N * nrhs * 16, hipMemcpyHostToDevice));
  })
}

int aifmm_ls_solve_resident(aifmm_ls_problem* h, double tol, int max_refine, double* stats) {
  GUARD({
    (void)tol;
    (void)max_refine;
    if (h->precision < 0) throw std::runtime_error("factor first");
    if (!h->rhs) throw std::runtime_error("set_rhs first");
    double t0 = now_s();
    if (h->precision == 0) h->f64->solve(h->rhs, h->sol, h->nrhs);
    else h->f32->solve(h->rhs, h->sol, h->nrhs);
    HIP_OK(hipDeviceSynchronize());
    h->t_solve = now_s() - t0;
    if (stats) {
      stats[0] = h->t_solve;
      stats[1] = 0;
      stats[2] = 0;
    }
  })
}

int aifmm_ls_get_solution(aifmm_ls_problem* h, double* x) {
  GUARD({
    if (!h->sol) throw std::runtime_error("no solution");
    HIP_OK(hipMemcpy(x, h->sol, h->pb->N * h->nrhs * 16, hipMemcpyDeviceToHost));
  })
}

int aifmm_ls_solve(aifmm_ls_problem* h, int nrhs, const double* b, double* x, double tol, int max_refine,
                   double* stats) {
  int rc = aifmm_ls_set_rhs(h, nrhs, b);
  if (rc) return rc;
  rc = aifmm_ls_solve_resident(h, tol, max_refine, stats);
  if (rc) return rc;
  return aifmm_ls_get_solution(h, x);
}

int aifmm_ls_stats(aifmm_ls_problem* h, double* out, int n) {
  GUARD({
    double v[12] = {0};
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
    }
    if (h->f64) v[7] = (double)h->f64->solve_bytes(1);
    if (h->f32) v[7] = (double)h->f32->solve_bytes(1);
    v[9] = (double)P.operator_bytes();
    v[10] = h->t_solve;
    v[11] = P.t_nca;
    for (int i = 0; i < n && i < 12; ++i) out[i] = v[i];
  })
}

}  // extern "C"
