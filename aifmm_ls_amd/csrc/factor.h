// Level-synchronous hierarchical elimination of the LS operator on the
// device, and the forward/base/backward solve sweeps over the stored factors.
#pragma once
#include <map>
#include <vector>

#include "problem.h"

namespace aifmm {

struct FactorStats {
  double t_factor = 0;     // seconds
  double flops = 0;        // algorithmic flops (LAPACK complex conventions)
  int base_n = 0;          // size of the dense level-2 system
  int max_rank = 0;
  size_t bytes = 0;        // factor storage
  int recompressions = 0;  // boxes re-based after fill-in
};

template <class T>
class Factor {
 public:
  explicit Factor(Problem& pb);
  void factor();
  // b, x: device FP64 (complex) arrays, N x nrhs column-major.
  void solve(const double2* b, double2* x, int nrhs);
  FactorStats st;
  size_t solve_bytes(int nrhs) const;  // algorithmic bytes of one sweep pair

  struct FBox {
    int m = 0, r = 0, color = 0;
    bool elim = false;
    Blk<T> L2P, P2M;
    std::map<int, Blk<T>> P2P, M2P, P2L, M2L;
    Blk<T> A;      // LU of the self block
    int* piv = nullptr;
    Blk<T> W;      // A^{-1} [L2P | P2P/M2P(k, neighbours)]
    Blk<T> G;      // first r columns: LU of L2M_f = -P2M A^{-1} L2P
    int* piv2 = nullptr;
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
  int* bpiv = nullptr;
  std::vector<int> boff;

  void init_level(int j);
  void nonleaf_p2p(int j);
  void eliminate_color(int j, int c);
  void recompress(int j);
  void base_factor();
  Blk<T> nb_blk(int j, int row, int col);  // coupling block (row eq, col unknown)
};

}  // namespace aifmm
