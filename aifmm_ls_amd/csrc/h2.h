// FP64 matrix-vector product with the assembled hierarchical operator
// A_H = P2P + L2P_L (M2L_L + L2P_{L-1} (M2L_{L-1} + ...) P2M_{L-1}) P2M_L,
// the operator the elimination inverts.  Used for iterative refinement of
// the complex64 factorisation and for residual checks.
#pragma once
#include "problem.h"

namespace aifmm {

class H2Op {
 public:
  explicit H2Op(Problem& pb) : pb(pb) {}
  // y = A_H x ; x, y device arrays N x nrhs (column-major, complex128)
  void apply(const double2* x, double2* y, int nrhs);
  // algorithmic bytes read per application (operator blocks + vectors)
  size_t bytes(int nrhs) const;

 private:
  Problem& pb;
  Arena ws{size_t(64) << 20};
  Arena sc{size_t(16) << 20};
};

// column 2-norms of an N x nrhs device array (host result)
void col_norms(Arena& sc, const double2* v, int64_t n, int nrhs, double* out, hipStream_t s);

}  // namespace aifmm
