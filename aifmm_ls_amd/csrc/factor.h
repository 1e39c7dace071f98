// Level-synchronous hierarchical elimination of the LS operator on the
// device, and the forward/base/backward solve sweeps over the stored factors.
#pragma once
#include <map>
#include <memory>
#include <type_traits>
#include <vector>

#include "problem.h"
#include "prog.h"

namespace aifmm {

struct FactorStats {
  double t_factor = 0;     // seconds
  double flops = 0;        // algorithmic flops (LAPACK complex conventions)
  int base_n = 0;          // size of the dense level-2 system
  int max_rank = 0;
  size_t bytes = 0;        // factor storage
  int recompressions = 0;  // boxes re-based after fill-in
};

// Factorisation of a square block for repeated solves: column-pivoted QR in
// complex128 (ColPivHouseholderQR, as the reference), LU with partial
// pivoting in complex64 (row operations keep the relative accuracy of the
// badly scaled rows that low-contrast boxes produce, which Householder
// reflections in float do not).
template <class T>
struct DFact {
  QRFact<T> q;  // q.A is the factored block in both cases
  int* piv = nullptr;
};

template <class T>
struct DenseOps {
  static constexpr bool LU = std::is_same<T, float2>::value;
  QRBatch<T> qf;
  QRSolveBatch<T> qs;
  LUBatch<T> lf;
  SolveBatch<T> ls;
  void factor(Arena& ar, Arena& sc, DFact<T>& f, const Blk<T>& A) {
    if (LU) {
      f.q.A = A;
      f.piv = ar.get<int>(A.r > 0 ? A.r : 1);
      lf.add(A, f.piv);
    } else {
      f.q = new_qr(ar, A);
      qf.add(f.q, sc);
    }
  }
  void solve(Arena& sc, const DFact<T>& f, const Blk<T>& B) {
    if (LU) ls.add(f.q.A, f.piv, B);
    else qs.add(f.q, B, sc);
  }
  void run_factor(Arena& sc, hipStream_t s) {
    qf.run(sc, s);
    lf.run(sc, s);
  }
  void run_solve(Arena& sc, hipStream_t s) {
    qs.run(sc, s);
    ls.run(sc, s);
  }
};

template <class T>
class Factor {
 public:
  explicit Factor(Problem& pb);
  void factor();
  // b, x: device FP64 (complex) arrays, N x nrhs column-major.  Launches are
  // asynchronous on the problem's stream (no host synchronisation).
  void solve(const double2* b, double2* x, int nrhs);
  FactorStats st;
  size_t solve_bytes(int nrhs) const;  // algorithmic bytes of one sweep pair

  struct FBox {
    int m = 0, r = 0, color = 0;
    bool elim = false;
    Blk<T> L2P, P2M;
    std::map<int, Blk<T>> P2P, M2P, P2L, M2L;
    Blk<T> A;      // self block, factored in place (DFact)
    DFact<T> fa;
    Blk<T> W;      // A^{-1} [L2P | P2P/M2P(k, neighbours)]
    Blk<T> G;      // first r columns: L2M_f = -P2M A^{-1} L2P, factored in place
    DFact<T> fg;
    // sweep operators: with W_r = W[:, :r] and L2M_f = G[:, :r],
    //   H = -L2M_f^{-1} P2M A^{-1}      (r x m)
    //   Z = A^{-1} - W_r H               (m x m)
    //   U = W_r L2M_f^{-1}               (m x r)
    Blk<T> Z, H, U;
  };
  std::vector<std::vector<FBox>> B;
  std::vector<std::vector<std::vector<int>>> colors;

 private:
  Problem& pb;
  int L, R;
  Arena ar{size_t(1) << 30};
  Arena sc{size_t(256) << 20};
  Arena sv{size_t(256) << 20};
  Blk<T> base;
  DFact<T> fb;
  Blk<T> binv;  // inverse of the level-2 system
  std::vector<int> boff;

  // sweep program and its vectors (capacity plan_cap right-hand sides)
  Arena pv{size_t(256) << 20};
  std::unique_ptr<Program<T>> plan;
  int plan_cap = 0;
  Blk<T> Bt, Xt, yb, xb;
  std::vector<std::vector<Blk<T>>> bk, bs, mr, xk, yk;
  void sweep_operators();
  void build_plan(int cap);

  void init_level(int j);
  void nonleaf_p2p(int j);
  void eliminate_color(int j, int c);
  void recompress(int j);
  void base_factor();
  Blk<T> nb_blk(int j, int row, int col);  // coupling block (row eq, col unknown)
};

}  // namespace aifmm
