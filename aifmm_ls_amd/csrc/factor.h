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
  int ill_boxes = 0;       // boxes whose sweeps keep the reference's association
  int dense_level = 0;     // top of the one-column sweep as one dense product: its level (0: staged)
  int dense_n = 0;         // and its size
  // one record per recompressed box, in order: level, colour step, box,
  // row-stack rank, column-stack rank, new rank
  std::vector<int> rc_trace;
  // complex128 only: every L2M_f factorisation whose nonzero-pivot count fell
  // below its size: level, box, nonzero pivots, size
  std::vector<int> zero_piv;
  double base_rdiag_ratio = 0;  // smallest over largest |R_ii| (complex128) or |U_ii| (complex64) of the base factor
  // complex64 with a growth gate: pivot blocks that kept their complex64 LU,
  // and those that went to FP64
  int piv_c64 = 0, piv_fp64 = 0;
};

// Factorisation of a square block for repeated solves.  complex128: column-
// pivoted QR, as the reference (ColPivHouseholderQR).  complex64: the block is
// widened to an FP64 copy and LU-factored in FP64 (the pivot blocks of the
// complex64 factorisation are the one place its rounding would be amplified
// by their conditioning); solves widen the right-hand sides, solve in FP64
// and round the result to complex64.
template <class T>
struct DFact {
  QRFact<T> q;                  // complex128: q.A is the factored block
  Blk<double2> a64;             // complex64: FP64 LU factors
  int* piv = nullptr;
  // complex64 with a growth gate (DenseOps<float2>::factor, gate_thr >= 0):
  // *flag = 0 when the block kept its complex64 LU (in q.A, in place), 1 when
  // it went to FP64 (a64); est = the block's growth-and-condition estimate
  int* flag = nullptr;
  float* est = nullptr;
};

template <class T>
struct DenseOps;

template <>
struct DenseOps<double2> {
  QRBatch<double2> qf;
  QRSolveBatch<double2> qs;
  void factor(Arena& ar, Arena& sc, DFact<double2>& f, const Blk<double2>& A, double nz_eps = 0,
              double gate_thr = -1) {
    (void)gate_thr;
    f.q = new_qr(ar, A);
    qf.add(f.q, sc, nz_eps);
  }
  void solve(Arena& sc, const DFact<double2>& f, const Blk<double2>& B) { qs.add(f.q, B, sc); }
  // the FP64 array holding the factor (R on and above the diagonal)
  static Blk<double2> fp64_factor(const DFact<double2>& f) { return f.q.A; }
  void run_factor(Arena& sc, hipStream_t s) { qf.run(sc, s); }
  void run_solve(Arena& sc, hipStream_t s) { qs.run(sc, s); }
};

// Complex64 pivot blocks.  Without a gate (gate_thr < 0) every block is
// widened and LU-factored in FP64 and its solves run in FP64.  With a gate
// (gate_thr >= 0; blocks of up to lu_lds_max(8) = 128 rows) each block is
// first LU-factored in complex64 in LDS and keeps that factor when its growth
// factor times its condition estimate stays at or below gate_thr; the others
// are factored in FP64, in the same launch when the FP64 factor fits the LDS
// (up to 100 rows), else by the FP64 kernels gated on flag 2.  The choice is
// made on the device, per block, and the solves follow it with no host round
// trip: one mixed solve launch, plus the gated FP64 kernels for blocks above
// 100 rows.
template <>
struct DenseOps<float2> {
  WidenBatch wf, ws;               // widen for factor / for solve
  CvtBatch<float2> nb;             // narrow solve results
  LUBatch<double2> lf;
  SolveBatch<double2> ls;
  GateBatch gf;                    // gated complex64 LUs (and their FP64 fallback)
  MixSolveBatch ms;                // solves with gated factors
  static Blk<double2> wide(Arena& a, const Blk<float2>& b) { return new_blk<double2>(a, b.r, b.c); }
  static void need_contig(const Blk<float2>& b) {
    if (b.size() && b.ld != b.r) throw std::runtime_error("FP64 pivot block path needs contiguous blocks");
  }
  // nz_eps applies to the QR path only: LU has no zero-pivot test
  void factor(Arena& ar, Arena& sc, DFact<float2>& f, const Blk<float2>& A, double nz_eps = 0,
              double gate_thr = -1) {
    (void)nz_eps;
    need_contig(A);
    f.q.A = A;
    f.a64 = wide(ar, A);
    f.piv = ar.get<int>(A.r > 0 ? A.r : 1);
    f.flag = nullptr;
    f.est = nullptr;
    if (!A.size()) return;
    if (gate_thr >= 0 && A.r <= lu_lds_max(sizeof(float2))) {
      f.flag = ar.get<int>(1);
      f.est = ar.get<float>(1);
      gf.d.push_back(GateDesc{A.p, f.piv, f.flag, f.a64.p, f.est, (float)gate_thr, A.r, A.ld, f.a64.ld});
      if (A.r > lu_lds_max(sizeof(double2))) {  // FP64 factor too large for the gate's LDS
        wf.add(A.p, f.a64.p, (int64_t)A.size(), f.flag, 2);
        lf.add(f.a64, f.piv, f.flag, 2);
      }
      return;
    }
    wf.add(A.p, f.a64.p, (int64_t)A.size());
    lf.add(f.a64, f.piv);
  }
  void solve(Arena& sc, const DFact<float2>& f, const Blk<float2>& B) {
    need_contig(B);
    if (!f.a64.r || !B.c) return;
    if (f.flag) {
      ms.d.push_back(MixSolveDesc{f.q.A.p, f.a64.p, f.piv, f.flag, B.p, f.a64.r, B.c, f.q.A.ld, f.a64.ld, B.ld});
      if (f.a64.r <= lu_lds_max(sizeof(double2))) return;
    }
    Blk<double2> b64 = wide(sc, B);
    const int want = f.flag ? 2 : 0;
    ws.add(B.p, b64.p, (int64_t)B.size(), f.flag, want);
    ls.add(f.a64, f.piv, b64, f.flag, want);
    nb.add(b64.p, B.p, (int64_t)B.size(), f.flag, want);
  }
  // the FP64 array holding the factor's diagonal (U on and above it in FP64;
  // a block that kept its complex64 LU has only its diagonal there)
  static Blk<double2> fp64_factor(const DFact<float2>& f) { return f.a64; }
  void run_factor(Arena& sc, hipStream_t s) {
    gf.run(sc, s);
    wf.run(sc, s);
    lf.run(sc, s);
  }
  void run_solve(Arena& sc, hipStream_t s) {
    ms.run(sc, s);
    ws.run(sc, s);
    ls.run(sc, s);
    nb.run(sc, s);
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
  size_t device_bytes() const;         // device memory held (arena chunks)
  // The pieces of solve() for callers that merge the sweeps of several
  // problems into one launch sequence (all on stream s): load b into the
  // sweep input, the sweep program (built for at least nrhs columns), store
  // the sweep output into x.
  Program<T>& sweep_program(int nrhs);
  // the sweep's most recent launch tuning (null before its first solve)
  const typename Program<T>::TuneResult* sweep_tune() const { return plan ? plan->last_tune() : nullptr; }
  void sweep_load(const double2* b, int nrhs, hipStream_t s);
  void sweep_store(double2* x, int nrhs, hipStream_t s);

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
    // A badly conditioned L2M_f (condition estimate above kIllCond) makes
    // Z and U large, and Z x0 - U y_k would cancel in rounding.  Such boxes
    // run the sweeps in the reference's order instead:
    //   forward   b_k = A^{-1} y - W_r (H y)
    //   backward  x_k = A^{-1} x0 - U (y_k - P2M A^{-1} x0)
    // so L2M_f^{-1} only meets the vectors the reference applies it to.
    bool ill = false;
    Blk<T> Ai;  // A^{-1}                (m x m)
    Blk<T> PA;  // -P2M A^{-1}           (r x m)
    // Composed sweep couplings (compose_sweep): for a regular neighbour nn,
    // P2P[nn] holds P2P[k][nn] Z_nn, and for a regular later-coloured k,
    // P2L[k] holds P2L[nn][k] Z_k; PU[nn] = P2P[k][nn] U_nn (original P2P)
    // for a regular later-coloured nn.
    std::map<int, Blk<T>> PU;
  };
  static constexpr double kIllCond = 1e4;
  // Zero-pivot epsilons of the complex128 factorisations (Eigen's
  // ColPivHouseholderQR "nonzero pivot" test with this epsilon in place of the
  // FP64 one; 0 = the FP64 epsilon, the reference's rule): the L2M_f pivot
  // blocks, and the level-2 base system.  Set before factor().
  double l2mf_eps = 1e-12, base_eps = 0;
  // complex64 only: growth gate of the pivot blocks (DenseOps<float2>;
  // negative: every pivot block in FP64).  Set before factor().
  double gate_thr = -1;
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
  // (vectors are complex128 in either precision: the box programs form every
  // product and sum in FP64)
  Blk<double2> Bt, Xt, yb, xb;
  std::vector<std::vector<Blk<double2>>> bk, bs, mr, xk, yk;
  std::vector<std::vector<Blk<double2>>> yt;  // phase-1 vectors of the back-substitution
  std::vector<std::vector<Blk<double2>>> vk;  // y_k - P2M A^{-1} x0 of the ill boxes
  // The other sweep program with its vectors: one-column sweeps keep their
  // own program (16-row tiles, the dense top) while wider ones use one built
  // for their width; swap_sweep exchanges the active set with this one.
  struct SweepSet {
    Arena pv{size_t(256) << 20};
    std::unique_ptr<Program<T>> plan;
    int plan_cap = 0;
    Blk<double2> Bt, Xt, yb, xb;
    std::vector<std::vector<Blk<double2>>> bk, bs, mr, xk, yk, yt, vk;
  } spare_;
  void swap_sweep();
  Blk<T> eye;                           // identity, rank x rank of the largest box
  // the sweep at levels jd..2 and the base solve as one dense matrix (n2 x
  // n2; n2 = 0: the staged form), level-jd box offsets in its stacked vectors
  static constexpr int kDenseTopMax = 6144;
  Blk<T> d2;
  int n2 = 0, jd = 2;
  std::vector<int> off2;
  void sweep_operators();
  void compose_sweep();
  void dense_top_alloc();
  void dense_top_fill();
  void alloc_vectors(Arena& a, int cap, int jhi, const Blk<double2>* b2, const Blk<double2>* x2);
  void emit_forward(Program<T>& P, int jlo, int jhi);
  void emit_base(Program<T>& P);
  void emit_backward(Program<T>& P, int jlo, int jhi);
  void build_plan(int cap);

  void init_level(int j);
  void nonleaf_p2p(int j);
  void eliminate_color(int j, int c);
  void recompress(int j, int c);
  void base_factor();
  Blk<T> nb_blk(int j, int row, int col);  // coupling block (row eq, col unknown)
  double t_wait_ = 0;                       // host seconds waiting at recompression readbacks
  double hs_[12] = {};                      // host seconds per sub-phase (AIFMM_FACTOR_TIMERS)
};

}  // namespace aifmm
