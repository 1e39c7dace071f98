// FP64 matrix-vector product with the assembled hierarchical operator
// A_H = P2P + L2P_L (M2L_L + L2P_{L-1} (M2L_{L-1} + ...) P2M_{L-1}) P2M_L,
// the operator the elimination inverts.  Used for iterative refinement of
// the complex64 factorisation and for residual checks.
#pragma once
#include "problem.h"
#include "prog.h"

#include <memory>

namespace aifmm {

class H2Op {
 public:
  explicit H2Op(Problem& pb) : pb(pb) {}
  // y = A_H x ; x, y device arrays N x nrhs (column-major, complex128).
  // Asynchronous on the problem's stream.  The launch program is built for
  // one (x, y) pair and rebuilt when either pointer or a larger nrhs appears.
  void apply(const double2* x, double2* y, int nrhs);
  // algorithmic bytes of one application (operator blocks + vectors)
  size_t bytes(int nrhs) const;

 private:
  Problem& pb;
  Arena ws{size_t(64) << 20};
  std::unique_ptr<Program<double2>> plan;
  const double2* px = nullptr;
  double2* py = nullptr;
  int cap = 0;
  void build(const double2* x, double2* y, int nrhs);
};

// column 2-norms of an N x nrhs device array (host result)
void col_norms(Arena& sc, const double2* v, int64_t n, int nrhs, double* out, hipStream_t s);

}  // namespace aifmm
