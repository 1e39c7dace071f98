// ORACLE — test infrastructure only.
//
// FP64 CPU restatement of the AIFMM_LS 2D Lippmann–Schwinger direct-solve
// path (reference: vaishna77/AIFMM_LS).  Only tests/, __graft_entry__.smoke()
// and bench.py's cpu_baseline leg may load this code, and only as the
// checker.  The product library (aifmm_ls_amd/) never links or calls it.
//
// Each function cites the reference file:line it restates.  Parity pins
// (what checks each part) live in tests/test_oracle_*.py; the functions
// that no check pins are marked "parity unpinned" where they are defined.
#pragma once
#include <complex>
#include <cstdint>
#include <map>
#include <vector>

namespace orc {

using cd = std::complex<double>;

// ---------------------------------------------------------------- dense
// Column-major complex matrix.  0x0 plays the role of Eigen's empty matrix
// (the reference tests `.size() == 0` on std::map<int,Mat> entries).
struct Mat {
  int r = 0, c = 0;
  std::vector<cd> a;
  Mat() {}
  Mat(int r_, int c_) : r(r_), c(c_), a((size_t)r_ * c_, cd(0, 0)) {}
  cd& operator()(int i, int j) { return a[(size_t)j * r + i]; }
  const cd& operator()(int i, int j) const { return a[(size_t)j * r + i]; }
  size_t size() const { return (size_t)r * c; }
};

Mat mul(const Mat& A, const Mat& B);
Mat mulAH(const Mat& A, const Mat& B);  // A^H * B
Mat adj(const Mat& A);
Mat neg(const Mat& A);
Mat block(const Mat& A, int i0, int j0, int nr, int nc);
Mat hcat(const Mat& A, const Mat& B);
Mat eye(int n);
double fro(const Mat& A);
void subeq(Mat& A, const Mat& B);  // A -= B (A 0x0 => A = -B)

// Partial-pivoted LU (reference uses Eigen::PartialPivLU for the skeleton
// transfer matrices, FMM2DTree_find_error.hpp:1118,1125).
struct LU {
  int n = 0;
  Mat f;
  std::vector<int> piv;
  void compute(const Mat& A);
  Mat solve(const Mat& B) const;
};

// Column-pivoted Householder QR with Eigen ColPivHouseholderQR semantics:
// norm downdating (LAPACK LAWN 176), rank() against a relative threshold,
// solve() on the nonzero pivots (reference calls at
// FMM2DTree_find_error.hpp:1535,1548,1564,1580,1819,2092,2477,2493,2510).
struct CPQR {
  int m = 0, n = 0;
  Mat qr;
  std::vector<cd> tau;
  std::vector<int> perm;  // perm[k] = original column placed at position k
  double maxpivot = 0;
  int nonzero = 0;
  // nz_eps: epsilon of Eigen's "nonzero pivot" test (0: the FP64 epsilon)
  void compute(const Mat& A, double nz_eps = 0);
  int rank(double threshold) const;
  Mat solve(const Mat& B) const;
  Mat Q(int k) const;  // first k columns of the unitary factor
};

// ---------------------------------------------------------------- math
void gauss_legendre(int n, std::vector<double>& x, std::vector<double>& w);
// J_n(x), Y_n(x) for n = 0..nmax (x > 0).  exact=true calls the C++17
// special functions per order (the reference calls Boost per term,
// LS_Assembly.hpp:39-65); exact=false uses Miller's backward recurrence for
// J (normalised by J0 + 2*sum J_2k = 1) and forward recurrence for Y.
void bessel_jy(double x, int nmax, double* J, double* Y, bool exact);

struct Params {
  int n_cheb = 6;
  int degree = 6;
  int n_levels = 5;
  int tol_pow = 8;
  int qchoice = 0;       // 0..5 as LS_functions.hpp; 6 = Gaussian-bump sum
  double kappa = 10.0;
  double L = 1.0;
  std::vector<double> bumps;  // a, w, cx, cy per bump (qchoice 6)
  int quantize = 1;           // round tables to 6 significant digits
  int exact_bessel = 0;       // per-term std:: special functions in entries
};

double contrast(const Params& P, double x, double y);

// Discretisation: points, contrast at points, far (M) and near (T) tables.
struct Disc {
  Params P;
  int nb_basis = 0;  // p(p+1)/2
  int rank = 0;      // n_cheb^2 points per leaf
  int n_leaf = 0;    // 4^L
  int64_t N = 0;
  double r_leaf = 0;
  std::vector<double> px, py, q;     // per point
  std::vector<double> cheb1d;        // standard Chebyshev nodes
  std::vector<int> basis_a, basis_b;
  Mat Qpinv;                         // nb_basis x rank
  Mat M;                             // 121 x rank (level L only)
  Mat T[9];                          // colleague operators at level L
  int series = 60;
  // Sensitivity probe (tests only): far-field entries are multiplied by
  // 1 + perturb * u(i, j), u a fixed hash of (i, j, perturb_seed) in [-1, 1].
  // Measures how far the algorithm's discrete choices (pivots, ranks) move
  // the solution under last-bit changes of its input.
  double perturb = 0;
  uint64_t perturb_seed = 0;
  // LS2DTree::findPhi (LS_Assembly.hpp:195, 237): true, the entries of the LS
  // operator I + kappa^2 q G; false, those of the volume potential G alone
  // (the scattered field from the solution, hybrid.cpp:264-281)
  bool find_phi = true;
  void build(const Params& P);
  void init_basis(const Params& P);  // nodes, basis, Q^+ only
  double box_radius(int j) const;
  // a precomputation table, unrounded: family -1 the far M at `level`
  // (121 x rank); 0..3 the colleague / fine / separated-fine / coarse
  // neighbour operator n of a target box at `level` (rank x rank)
  Mat table(int family, int level, int n) const;
  cd entry(int64_t i, int64_t j) const;
  cd far_sum(int64_t i, int leaf_j, int idx_j) const;
};

// The 6-significant-digit rounding of the precomputation text cache.
cd round6c(cd v);
// The Graf series of a far entry: sum_n M(n + series, idx_j) H_n(kappa rho)
// e^{i n phi}, (rho, phi) the polar form of (x - cx, y - cy).
cd graf_series(const Mat& M, int idx_j, double x, double y, double cx, double cy, double kappa, int series,
               bool exact);

// Adaptive quadtree with 2:1 level restriction and the entries between its
// leaves (LS2DTree::createAdaptiveTree, makeLevelRestriction_Tree,
// getIfNeighbor, getMatrixEntry; LS_Assembly.hpp:283-541, 609-742, 892-1072).
// Laid out as the reference keeps it: per level the boxes in creation order,
// indexTree holding their box numbers, leafNodes the (level, index) pairs.
struct AdaptiveDisc {
  struct ABox {
    double cx = 0, cy = 0, radius = 0;
    int level = 0, boxNumber = -1;
    bool isLeaf = false;
  };
  Disc D;  // nodes, basis, Q^+ and the table quadrature
  double epsilon_tree = 0.1;
  int n_levels = 0, rank = 0;
  int restriction_splits = 0;  // boxes refineBox split
  std::vector<std::vector<ABox>> tree;
  std::vector<std::vector<int>> indexTree;
  std::vector<std::pair<int, int>> leafNodes;
  std::vector<double> boxRadius;
  std::vector<double> px, py, q;
  int64_t N = 0;
  bool find_phi = true;
  void build(const Params& P, int tree_adaptivity);
  // getIfNeighbor between leaves li (target) and lj: the reference's
  // typeOfNeighbor (-1 far, 0 colleague, 1 coarse, 2 fine, 3 separated fine)
  // and NumOfNeighbor (-1 when no slot matched)
  int neighbor(int li, int lj, int& num) const;
  cd entry(int64_t i, int64_t j);
  const Mat& near_table(int type, int level, int num);
  const Mat& far_table(int level);

 private:
  std::map<std::vector<int>, Mat> ntab_;  // (type, level, num) -> table, rounded as the cache holds it
  std::map<int, Mat> ftab_;               // level -> M
};

// Quadtree index helpers (reference child order 0=(-,-) 1=(+,-) 2=(+,+)
// 3=(-,+), LS_Assembly.hpp:834-849).
void box_xy(int level, int b, int& ix, int& iy);
int xy_box(int level, int ix, int iy);
extern const int NB_OFF[8][2];
extern const int INNER_OFF[16][2];
extern const int OUTER_OFF[24][2];

enum Schedule { SCHED_EFFICIENT = 0, SCHED_EAGER = 1, SCHED_COLORED = 2 };

struct Box {
  int nb[8], inner[16], outer[24];
  bool Eliminated = false;
  int NumMultipoles = 0, NumLocals = 0;
  std::vector<int> charges;  // leaf point indices
  std::vector<int> in_chk, in_chg, out_chk, out_chg;
  Mat Ac, Ar, L2P, P2M, L2P_f, L2M_f;
  std::map<int, Mat> M2L, P2P, M2P, P2L, P2M_f, M2M_f;
  Mat particle_rhs, multipole_rhs, local_rhs, particles, multipoles, locals;
};

struct SolveStats {
  double t_assemble = 0, t_factor = 0, t_solve = 0;
  int max_rank[16] = {0};
  double mean_rank[16] = {0};
};

// FMM2DTree restatement (FMM2DTree_find_error.hpp:76-2583).
struct Tree {
  const Disc* D = nullptr;
  int L = 0;
  int tol_pow = 8;
  double rrqr_thr = 1e-8;
  // Zero-pivot epsilon of the L2M_f factorisations (eliminate_z and the
  // back-substitution's Atemp).  0 keeps Eigen's rule (the FP64 epsilon), as
  // the reference; a positive value counts pivots below nz_eps times the
  // largest column norm as zero, the rule of the GPU's complex128 factor.
  double piv_eps = 0;
  // The same rule for the level-2 base system (0: Eigen's rule, the reference's).
  double base_piv_eps = 0;
  // Every L2M_f factorisation of the elimination whose nonzero-pivot count
  // fell below its size: level, box, nonzero pivots, size.
  std::vector<int> zero_piv;
  // smallest over largest |R_ii| of the base system's factorisation (0 if empty)
  double base_rdiag_ratio = 0;
  // Colour schedule only: one record per recompressed box, in order:
  // level, colour step, box, row-stack rank, column-stack rank, new rank.
  std::vector<int> rc_trace;
  int cur_color = -1;
  std::vector<std::vector<Box>> T;
  std::vector<std::vector<int>> order;  // elimination order per level
  std::vector<std::pair<int, int>> P2Plist, P2L_M2Plist;
  void create(const Disc* D, int tol_pow);
  void assign_rhs(const cd* b);
  void getNodes();
  void assemble_M2L();
  void initialise_phase();
  void initialise_P2P_leaf();
  void initialise_P2P_nonleaf(int j);
  void eliminate(Schedule s);
  void base_solve();
  void back_substitution();
  void solution(cd* x) const;
  // pieces
  void aca(const std::vector<int>& rows, const std::vector<int>& cols,
           std::vector<int>& rb, std::vector<int>& cb, Mat& Ac, Mat& Ar) const;
  void nodes_outgoing(int j, int k);
  void nodes_incoming(int j, int k);
  void l2p_p2m(int j, int k);
  void eliminate_x(int j, int k);
  void eliminate_z(int j, int k);
  bool well_separated(int j, int k, int l) const;
  void compress_P2P_two_ways(int j, int k, int b);
  void compress_M2P_P2L(int j, int k, int b);
  void compress_P2P(int j, int k, int b, Mat& nL2P, Mat& nP2Mt, Mat& M2L);
  void compress_M2P(int j, int k, int b, Mat& nL2P, Mat& M2L);
  void compress_P2L(int j, int k, int b, Mat& nP2Mt, Mat& M2L);
  void update_M2L_at_k(int j, int k, int b, const Mat& nL2P);
  void update_M2L_at_b(int j, int k, int b, const Mat& nP2Mt);
  void update_parent_L2P(int j, int k, const Mat& nL2P);
  void update_parent_P2M(int j, int b, const Mat& nP2Mt);
  void colored_recompress(int j);
};

}  // namespace orc
