// One discretised Lippmann–Schwinger problem on the device: points and
// contrast, the precomputed tables, the quadtree interaction lists and the
// FP64 hierarchical operator (NCA skeletons, L2P/P2M, M2L, leaf P2P).
#pragma once
#include <map>
#include <vector>

#include "batch.h"

namespace aifmm {

struct Params {
  int n_cheb = 6, degree = 6, n_levels = 5, tol_pow = 8, qchoice = 0, quantize = 1;
  double kappa = 10.0, L = 1.0;
  std::vector<double> bumps;  // a, w, cx, cy per bump (qchoice 6)
};

struct Lists {
  int nb[8], inner[16], outer[24];
};

struct SkelBox {
  std::vector<int64_t> in_chk, in_chg, out_chk, out_chg;
  int r = 0;  // skeleton rank (incoming == outgoing after truncation)
  int m = 0;  // rows of L2P: points (leaf) or sum of children ranks
  Blk<double2> L2P, P2M;
  std::map<int, Blk<double2>> M2L;  // interaction-list blocks
};

double contrast(const Params& P, double x, double y);
void gauss_legendre10(std::vector<double>& x, std::vector<double>& w);

class Problem {
 public:
  explicit Problem(const Params& P);
  ~Problem();
  Params P;
  int R = 36, nbasis = 21, L = 5, nleaf = 0;
  int64_t N = 0;
  double rleaf = 0;
  std::vector<double> px, py, q;
  std::vector<double2> M_h, T_h;  // rounded tables (host copies)
  double *dpx = nullptr, *dpy = nullptr, *dq = nullptr;
  double2 *dM = nullptr, *dT = nullptr;
  int64_t* d_iota = nullptr;  // 0..N-1 (leaf charge lists)
  EntryCtx ctx{};
  std::vector<std::vector<Lists>> lists;
  std::vector<std::vector<SkelBox>> sk;
  std::vector<std::map<int, Blk<double2>>> P2P;  // leaf near-field blocks
  bool assembled = false;
  int quad_capped = 0;
  double t_tables = 0, t_nca = 0, t_assemble = 0;
  Arena persist{size_t(64) << 20};
  Arena op{size_t(1) << 30};
  Arena scratch{size_t(256) << 20};
  Arena acaws{size_t(1) << 30};
  Arena acakeep{size_t(256) << 20};
  hipStream_t stream = nullptr;

  void build_tables();
  void build_lists();
  void assemble();
  // LS2DTree::findPhi of entries() and the dense product (LS_Assembly.hpp:195,
  // 507-512, 530-535): 1, the LS operator I + kappa^2 q G; 0, the volume
  // potential G alone.  Assembly and factorisation always use the LS operator.
  int find_phi = 1;
  void entries(int64_t n, const int64_t* I, const int64_t* J, double2* out);
  // the entry context of entries() and the dense product (find_phi applied)
  EntryCtx ctx_phi() const {
    EntryCtx c = ctx;
    c.find_phi = find_phi;
    return c;
  }
  int well_sep(int j, int k, int l) const;
  size_t operator_bytes() const { return op.used(); }
  size_t device_bytes() const;  // device memory held (arena chunks)

 private:
  void nca_level(int j);
};

}  // namespace aifmm
