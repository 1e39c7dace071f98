// ORACLE — test infrastructure only (see ls_oracle.hpp).
// C entry points used by the Python tests and by bench.py's cpu_baseline.
#include <chrono>
#include <cmath>
#include <cstring>
#include <stdexcept>

#include "ls_oracle.hpp"

using namespace orc;

namespace {
struct Handle {
  Disc D;
  Tree T;
  std::vector<cd> b, x;
  double t_tables = 0, t_assemble = 0, t_factor = 0, t_solve = 0;
  double piv_eps = 0;       // Tree::piv_eps of every later solve
  double base_piv_eps = 0;  // Tree::base_piv_eps of every later solve
  bool assembled = false;
};
double now() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}
}  // namespace

extern "C" {

// params: n_cheb, degree, n_levels, tol_pow, qchoice, quantize, exact_bessel
void* orc_create(const int* ip, double kappa, double L, const double* bumps, int nbumps) {
  Handle* h = new Handle;
  Params P;
  P.n_cheb = ip[0];
  P.degree = ip[1];
  P.n_levels = ip[2];
  P.tol_pow = ip[3];
  P.qchoice = ip[4];
  P.quantize = ip[5];
  P.exact_bessel = ip[6];
  P.kappa = kappa;
  P.L = L;
  if (bumps && nbumps > 0) P.bumps.assign(bumps, bumps + 4 * nbumps);
  double t0 = now();
  h->D.build(P);
  h->t_tables = now() - t0;
  return h;
}

void orc_destroy(void* p) { delete (Handle*)p; }

// Sensitivity probe: perturb the far-field entries (see Disc::perturb).
void orc_set_perturb(void* p, double eps, unsigned long long seed) {
  Handle* h = (Handle*)p;
  h->D.perturb = eps;
  h->D.perturb_seed = seed;
}

// Zero-pivot epsilon of the L2M_f factorisations in later solves (Tree::piv_eps;
// 0 = Eigen's rule, the reference's).
void orc_set_piv_eps(void* p, double eps, double base_eps) {
  ((Handle*)p)->piv_eps = eps;
  ((Handle*)p)->base_piv_eps = base_eps;
}

// L2M_f factorisations of the last solve with zeroed pivots (Tree::zero_piv,
// 4 ints each); writes up to cap ints, returns the number of ints.
int orc_zero_pivots(void* p, int* out, int cap) {
  const std::vector<int>& t = ((Handle*)p)->T.zero_piv;
  for (int i = 0; i < (int)t.size() && i < cap; ++i) out[i] = t[i];
  return (int)t.size();
}

// Recompression records of the last colour-schedule solve (Tree::rc_trace,
// 6 ints each); writes up to cap ints, returns the number of ints.
int orc_recompress_trace(void* p, int* out, int cap) {
  Handle* h = (Handle*)p;
  const std::vector<int>& t = h->T.rc_trace;
  for (int i = 0; i < (int)t.size() && i < cap; ++i) out[i] = t[i];
  return (int)t.size();
}

// LS2DTree::findPhi of later entries and dense matrices (Disc::find_phi).
void orc_set_find_phi(void* p, int on) { ((Handle*)p)->D.find_phi = on != 0; }

long long orc_n(void* p) { return ((Handle*)p)->D.N; }

void orc_points(void* p, double* xy, double* q) {
  Handle* h = (Handle*)p;
  for (int64_t i = 0; i < h->D.N; ++i) {
    xy[2 * i] = h->D.px[i];
    xy[2 * i + 1] = h->D.py[i];
    if (q) q[i] = h->D.q[i];
  }
}

// M: 121 x rank (column-major, interleaved complex); T: 9 x rank x rank
void orc_tables(void* p, double* M, double* T, double* Qpinv) {
  Handle* h = (Handle*)p;
  if (M) std::memcpy(M, h->D.M.a.data(), h->D.M.a.size() * sizeof(cd));
  if (T)
    for (int n = 0; n < 9; ++n)
      std::memcpy(T + (size_t)n * h->D.rank * h->D.rank * 2, h->D.T[n].a.data(), h->D.T[n].a.size() * sizeof(cd));
  if (Qpinv) std::memcpy(Qpinv, h->D.Qpinv.a.data(), h->D.Qpinv.a.size() * sizeof(cd));
}

// One precomputation table, unrounded (Disc::table): family -1 the far M at
// `level` (121 x rank), 0..3 a neighbour operator (rank x rank); column-major
// interleaved complex.  Returns the number of rows, or -1 for no such table.
int orc_table(void* p, int family, int level, int n, double* out) {
  Handle* h = (Handle*)p;
  try {
    Mat t = h->D.table(family, level, n);
    std::memcpy(out, t.a.data(), t.a.size() * sizeof(cd));
    return t.r;
  } catch (const std::exception&) {
    return -1;
  }
}

void orc_entries(void* p, long long n, const long long* I, const long long* J, double* out) {
  Handle* h = (Handle*)p;
#pragma omp parallel for schedule(static)
  for (long long t = 0; t < n; ++t) {
    cd v = h->D.entry(I[t], J[t]);
    out[2 * t] = v.real();
    out[2 * t + 1] = v.imag();
  }
}

void orc_bessel(double x, int nmax, int exact, double* J, double* Y) { bessel_jy(x, nmax, J, Y, exact != 0); }

void orc_gauss_legendre(int n, double* x, double* w) {
  std::vector<double> xs, ws;
  gauss_legendre(n, xs, ws);
  for (int i = 0; i < n; ++i) x[i] = xs[i], w[i] = ws[i];
}

// Plane-wave RHS b_i = -kappa^2 q(x_i) exp(i kappa d.x_i), d = (cos th, sin th)
// (LS_functions.hpp:113-120 for th = 0).
void orc_rhs(void* p, double theta, double* b) {
  Handle* h = (Handle*)p;
  double k = h->D.P.kappa;
  for (int64_t i = 0; i < h->D.N; ++i) {
    double ph = k * (cos(theta) * h->D.px[i] + sin(theta) * h->D.py[i]);
    cd v = -1.0 * k * k * h->D.q[i] * std::exp(cd(0.0, 1.0) * ph);
    if (theta == 0.0) v = -1.0 * k * k * h->D.q[i] * std::exp(cd(0.0, 1.0) * k * h->D.px[i]);
    b[2 * i] = v.real();
    b[2 * i + 1] = v.imag();
  }
}

// Full hierarchical solve: NCA, M2L/P2P assembly, elimination with the given
// schedule, base solve, back-substitution.  stats (optional, >= 8 doubles):
// t_tables, t_assemble, t_factor, t_solve, sum of leaf ranks, max rank,
// base-system size, smallest over largest |R_ii| of the base factorisation.
int orc_solve(void* p, const double* b, double* x, int schedule, double* stats) {
  Handle* h = (Handle*)p;
  int64_t N = h->D.N;
  std::vector<cd> bb(N), xx(N);
  for (int64_t i = 0; i < N; ++i) bb[i] = cd(b[2 * i], b[2 * i + 1]);
  Tree& T = h->T;
  T = Tree();
  T.piv_eps = h->piv_eps;
  T.base_piv_eps = h->base_piv_eps;
  double t0 = now();
  T.create(&h->D, h->D.P.tol_pow);
  T.assign_rhs(bb.data());
  T.getNodes();
  T.assemble_M2L();
  T.initialise_phase();
  T.initialise_P2P_leaf();
  double t1 = now();
  T.eliminate((Schedule)schedule);
  double t2 = now();
  T.back_substitution();
  double t3 = now();
  T.solution(xx.data());
  for (int64_t i = 0; i < N; ++i) x[2 * i] = xx[i].real(), x[2 * i + 1] = xx[i].imag();
  if (stats) {
    stats[0] = h->t_tables;
    stats[1] = t1 - t0;
    stats[2] = t2 - t1;
    stats[3] = t3 - t2;
    double s = 0, mx = 0;
    for (auto& B : T.T[T.L]) s += B.in_chk.size(), mx = std::max(mx, (double)B.in_chk.size());
    stats[4] = s;
    stats[5] = mx;
    int nb = 0;
    for (auto& B : T.T[2]) nb += B.NumMultipoles;
    stats[6] = nb;
    stats[7] = T.base_rdiag_ratio;
  }
  return 0;
}

// Per-level skeleton ranks after NCA truncation (in_chk sizes), level-major:
// out[j*4^L + k] for j = 2..L (boxes beyond 4^j are left untouched).
void orc_ranks(void* p, int* out) {
  Handle* h = (Handle*)p;
  int L = h->T.L;
  int nl = 1 << (2 * L);
  for (int j = 2; j <= L; ++j)
    for (int k = 0; k < (1 << (2 * j)); ++k) out[(size_t)j * nl + k] = (int)h->T.T[j][k].in_chk.size();
}

// Skeleton index lists of one box after NCA (for tests).
int orc_skeleton(void* p, int j, int k, int which, int* out) {
  Handle* h = (Handle*)p;
  const Box& B = h->T.T[j][k];
  const std::vector<int>* v = which == 0 ? &B.in_chk : which == 1 ? &B.in_chg : which == 2 ? &B.out_chk : &B.out_chg;
  for (size_t i = 0; i < v->size(); ++i) out[i] = (*v)[i];
  return (int)v->size();
}

// Run only the NCA + transfer operators (no factorisation).
void orc_nca(void* p) {
  Handle* h = (Handle*)p;
  h->T = Tree();
  h->T.create(&h->D, h->D.P.tol_pow);
  h->T.getNodes();
}

// One ACA call on explicit index lists (rows x cols); returns rank, writes
// pivot rows/cols and the un-updated pivot columns Ac (N1 x rank) and rows
// Ar (rank x N2), column-major interleaved complex.
int orc_aca(void* p, int tol_pow, int n1, const int* rows, int n2, const int* cols, int* rb, int* cb, double* Ac,
            double* Ar) {
  Handle* h = (Handle*)p;
  Tree t;
  t.D = &h->D;
  t.tol_pow = tol_pow;
  std::vector<int> R(rows, rows + n1), C(cols, cols + n2), vb, cbv;
  Mat A1, A2;
  t.aca(R, C, vb, cbv, A1, A2);
  for (size_t i = 0; i < vb.size(); ++i) rb[i] = vb[i], cb[i] = cbv[i];
  if (Ac) std::memcpy(Ac, A1.a.data(), A1.a.size() * sizeof(cd));
  if (Ar) std::memcpy(Ar, A2.a.data(), A2.a.size() * sizeof(cd));
  return (int)vb.size();
}

// Dense matrix A (N x N, column-major, interleaved complex); small N only.
void orc_dense(void* p, double* A) {
  Handle* h = (Handle*)p;
  int64_t N = h->D.N;
#pragma omp parallel for schedule(dynamic)
  for (int64_t j = 0; j < N; ++j)
    for (int64_t i = 0; i < N; ++i) {
      cd v = h->D.entry(i, j);
      A[2 * (j * N + i)] = v.real();
      A[2 * (j * N + i) + 1] = v.imag();
    }
}

// Test hooks for the dense kernels: CPQR of A (m x n) -> rank at threshold,
// Q (m x k) and R-diagonal magnitudes; LU solve.
int orc_cpqr(int m, int n, const double* A, double thr, int k, double* Q, double* rdiag, int* perm) {
  Mat M(m, n);
  std::memcpy(M.a.data(), A, sizeof(cd) * m * n);
  CPQR q;
  q.compute(M);
  if (Q) {
    Mat QQ = q.Q(k);
    std::memcpy(Q, QQ.a.data(), sizeof(cd) * m * k);
  }
  if (rdiag)
    for (int i = 0; i < std::min(m, n); ++i) rdiag[i] = std::abs(q.qr(i, i));
  if (perm)
    for (int i = 0; i < n; ++i) perm[i] = q.perm[i];
  return q.rank(thr);
}

void orc_qr_solve(int n, const double* A, int nrhs, const double* B, double* X) {
  Mat M(n, n), R(n, nrhs);
  std::memcpy(M.a.data(), A, sizeof(cd) * n * n);
  std::memcpy(R.a.data(), B, sizeof(cd) * n * nrhs);
  CPQR q;
  q.compute(M);
  Mat S = q.solve(R);
  std::memcpy(X, S.a.data(), sizeof(cd) * n * nrhs);
}

void orc_lu_solve(int n, const double* A, int nrhs, const double* B, double* X) {
  Mat M(n, n), R(n, nrhs);
  std::memcpy(M.a.data(), A, sizeof(cd) * n * n);
  std::memcpy(R.a.data(), B, sizeof(cd) * n * nrhs);
  LU lu;
  lu.compute(M);
  Mat S = lu.solve(R);
  std::memcpy(X, S.a.data(), sizeof(cd) * n * nrhs);
}

}  // extern "C"

extern "C" {
// Interaction lists of one box (8 neighbour, 16 inner, 24 outer slots; -1 = none).
void orc_lists(void* p, int level, int box, int* out) {
  Handle* h = (Handle*)p;
  Tree t;
  t.create(&h->D, h->D.P.tol_pow);
  const Box& B = t.T[level][box];
  for (int s = 0; s < 8; ++s) out[s] = B.nb[s];
  for (int s = 0; s < 16; ++s) out[8 + s] = B.inner[s];
  for (int s = 0; s < 24; ++s) out[24 + s] = B.outer[s];
}
}

// ---------------------------------------------------------- adaptive tree
extern "C" {
// params as orc_create; tree_adaptivity as LS2DTree's (epsilon 10^-t)
void* orc_adaptive_create(const int* ip, double kappa, double L, const double* bumps, int nbumps,
                          int tree_adaptivity) {
  Params P;
  P.n_cheb = ip[0];
  P.degree = ip[1];
  P.n_levels = ip[2];
  P.tol_pow = ip[3];
  P.qchoice = ip[4];
  P.quantize = ip[5];
  P.exact_bessel = ip[6];
  P.kappa = kappa;
  P.L = L;
  if (bumps && nbumps > 0) P.bumps.assign(bumps, bumps + 4 * nbumps);
  AdaptiveDisc* a = new AdaptiveDisc;
  a->build(P, tree_adaptivity);
  return a;
}
void orc_adaptive_destroy(void* p) { delete (AdaptiveDisc*)p; }
long long orc_adaptive_n(void* p) { return ((AdaptiveDisc*)p)->N; }
int orc_adaptive_depth(void* p) { return ((AdaptiveDisc*)p)->n_levels; }
int orc_adaptive_splits(void* p) { return ((AdaptiveDisc*)p)->restriction_splits; }
// leaves in leafNodes order: level, box number, centre (x, y); returns the count
int orc_adaptive_leaves(void* p, int* levels, long long* boxes, double* centres, int cap) {
  AdaptiveDisc* a = (AdaptiveDisc*)p;
  const int n = (int)a->leafNodes.size();
  for (int k = 0; k < n && k < cap; ++k) {
    const auto& b = a->tree[a->leafNodes[k].first][a->leafNodes[k].second];
    levels[k] = b.level;
    boxes[k] = b.boxNumber;
    centres[2 * k] = b.cx;
    centres[2 * k + 1] = b.cy;
  }
  return n;
}
void orc_adaptive_points(void* p, double* xy, double* q) {
  AdaptiveDisc* a = (AdaptiveDisc*)p;
  for (int64_t i = 0; i < a->N; ++i) xy[2 * i] = a->px[i], xy[2 * i + 1] = a->py[i], q[i] = a->q[i];
}
// getIfNeighbor of leaf pairs: type (-1 far, 0 colleague, 1 coarse, 2 fine,
// 3 separated fine) and slot
void orc_adaptive_neighbor(void* p, int n, const int* li, const int* lj, int* type, int* num) {
  AdaptiveDisc* a = (AdaptiveDisc*)p;
  for (int t = 0; t < n; ++t) type[t] = a->neighbor(li[t], lj[t], num[t]);
}
void orc_adaptive_set_find_phi(void* p, int on) { ((AdaptiveDisc*)p)->find_phi = on != 0; }
int orc_adaptive_entries(void* p, long long n, const long long* I, const long long* J, double* out) {
  AdaptiveDisc* a = (AdaptiveDisc*)p;
  try {
    for (long long t = 0; t < n; ++t) {
      cd e = a->entry(I[t], J[t]);
      out[2 * t] = e.real();
      out[2 * t + 1] = e.imag();
    }
  } catch (const std::exception&) {
    return -1;
  }
  return 0;
}
// y = A x over every entry (small N only)
int orc_adaptive_dense_matvec(void* p, const double* x, double* y) {
  AdaptiveDisc* a = (AdaptiveDisc*)p;
  const int64_t N = a->N;
  try {
    // tables first, serially (the caches are not thread-safe)
    for (int64_t i = 0; i < N; i += a->rank)
      for (int64_t j = 0; j < N; j += a->rank) a->entry(i, j);
  } catch (const std::exception&) {
    return -1;
  }
#pragma omp parallel for schedule(dynamic, 16)
  for (int64_t i = 0; i < N; ++i) {
    cd s = 0;
    for (int64_t j = 0; j < N; ++j) s += a->entry(i, j) * cd(x[2 * j], x[2 * j + 1]);
    y[2 * i] = s.real();
    y[2 * i + 1] = s.imag();
  }
  return 0;
}
}
