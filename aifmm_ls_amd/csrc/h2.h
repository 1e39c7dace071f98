// FP64 matrix-vector product with the assembled hierarchical operator
// A_H = P2P + L2P_L (M2L_L + L2P_{L-1} (M2L_{L-1} + ...) P2M_{L-1}) P2M_L,
// the operator the elimination inverts.  Used for iterative refinement of
// the complex64 factorisation and for residual checks.
#pragma once
#include "problem.h"
#include "prog.h"
#include "kernels.h"

#include <memory>

namespace aifmm {

class H2Op {
 public:
  explicit H2Op(Problem& pb) : pb(pb) { ws.set_chunk(Arena::chunk_for(pb.N, 512, 16 << 20, 64 << 20)); }
  // y = A_H x ; x, y device arrays N x nrhs (column-major, complex128).
  // Asynchronous on the problem's stream.  The launch program works on its
  // own input and output buffers (x and y are copied in and out), built once
  // per right-hand-side capacity.
  void apply(const double2* x, double2* y, int nrhs);
  // algorithmic bytes of one application (operator blocks + vectors)
  size_t bytes(int nrhs) const;
  size_t device_bytes() const { return ws.reserved(); }
  // The pieces of apply() for callers that merge several problems' launch
  // sequences (all on stream s): the program (built for nrhs columns), load
  // x into its input, store its output into y.
  Program<double2>& program(int nrhs);
  void load(const double2* x, int nrhs, hipStream_t s);
  void store(double2* y, int nrhs, hipStream_t s);
  // The near field, added to the program's output after it ran: launched
  // here (near), or described for a launch over several problems.
  void near(int nrhs, hipStream_t s);
  NearDesc near_desc(int nrhs) const;

 private:
  Problem& pb;
  Arena ws{size_t(64) << 20};
  std::unique_ptr<Program<double2>> plan;
  double2* xin = nullptr;
  double2* yout = nullptr;
  int* dnbr = nullptr;      // leaf boxes x 9 near-field neighbours
  double2* dtk = nullptr;   // kappa^2 x the near-field tables
  Arena stage_dev{size_t(1) << 20};  // the near-field descriptor
  int cap = 0;
  void build(int nrhs);
};

}  // namespace aifmm
