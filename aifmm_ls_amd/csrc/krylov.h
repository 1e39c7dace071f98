// FP64 refinement of a solve against the hierarchical operator A_H.
//
// Flexible GMRES(m), right-preconditioned by the factorisation's sweep M
// (z = M r is a forward sweep + base solve + back-substitution): the
// preconditioned vectors z_j are kept, so M may be the complex64 sweep,
// whose rounding makes it only approximately linear.  Arnoldi uses classical
// Gram–Schmidt with one re-orthogonalisation; every right-hand side has its
// own Arnoldi process, run in lockstep on N x nrhs blocks.  Plain Richardson
// refinement (x += M (b - A x)) is kept as the alternative.
#pragma once
#include <functional>
#include <vector>

#include "h2.h"

namespace aifmm {

struct RefineResult {
  int steps = 0;       // preconditioner applications after the first
  double relres = -1;  // max over columns of |b - A x| / |b| at exit
};

class Refiner {
 public:
  using Precond = std::function<void(const double2* r, double2* z)>;  // N x nrhs each
  Refiner(Problem& pb, H2Op& A) : pb_(pb), A_(A) { ws_.set_chunk(Arena::chunk_for(pb.N, 1 << 10, 16 << 20, 256 << 20)); }
  size_t device_bytes() const { return ws_.reserved(); }
  // x <- solution of A_H x = b to relative residual tol (at most maxit
  // preconditioner applications after x0 = M b, or after the x passed in when
  // x_given).  method 0: FGMRES, 1: Richardson.
  RefineResult run(const Precond& M, const double2* b, double2* x, int nrhs, double tol, int maxit, int method,
                   bool x_given = false);

 private:
  Problem& pb_;
  H2Op& A_;
  Arena ws_{size_t(256) << 20};
  double2* tmp_ = nullptr;
  size_t cap_ = 0;
  // per-column 2-norms of v (N x nrhs) via partial sums summed on the host
  void norms(const double2* v, int nrhs, std::vector<double>& out);
  void dots(const double2* V, int64_t vstride, int nv, const double2* w, int nrhs, std::vector<double2>& out);
  void lincomb(double2* w, const double2* V, int64_t vstride, int nv, const std::vector<double2>& coef,
               const std::vector<double2>& beta, int nrhs);
  double2* dpart_ = nullptr;
  double2* dcoef_ = nullptr;
  double2* dbeta_ = nullptr;
};

// Flexible GMRES(m) for a batch of independent problems in lockstep, one
// right-hand side each: every step applies all members' preconditioners in
// one merged launch sequence and all their operators in another (the
// callbacks), and the members' dot products come back to the host together.
// Each member has its own Arnoldi process; members that have converged keep
// stepping until all have (their corrections are then below the tolerance).
struct BatchVec {
  int64_t N;
  const double2* b;
  double2* x;
};

class BatchRefiner {
 public:
  // dst[i] = Op_i(src[i]) for every member, asynchronous on stream s
  using Apply = std::function<void(const std::vector<const double2*>& src, const std::vector<double2*>& dst)>;
  // x_i <- solution of A_i x = b_i to relative residual tol, starting from
  // x_i = M_i b_i; at most maxit preconditioner applications after that.
  std::vector<RefineResult> run(const std::vector<BatchVec>& v, const Apply& M, const Apply& A, double tol, int maxit,
                                hipStream_t s);

 private:
  Arena ws_{size_t(256) << 20};
};

}  // namespace aifmm
