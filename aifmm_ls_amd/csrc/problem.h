// One discretised Lippmann–Schwinger problem on the device: points and
// contrast, the precomputed tables, the quadtree interaction lists and the
// FP64 hierarchical operator (NCA skeletons, L2P/P2M, M2L, leaf P2P).
#pragma once
#include <map>
#include <string>
#include <vector>

#include "batch.h"

namespace aifmm {

struct Params {
  int n_cheb = 6, degree = 6, n_levels = 5, tol_pow = 8, qchoice = 0, quantize = 1;
  double kappa = 10.0, L = 1.0;
  std::vector<double> bumps;  // a, w, cx, cy per bump (qchoice 6)
  // adaptive = 1: the adaptive quadtree with 2:1 level restriction in place of
  // the uniform one (n_levels unused), refined until the contrast is
  // interpolated to 10^-tree_adaptivity (LS_Assembly.hpp:201, 628-742)
  int adaptive = 0, tree_adaptivity = 1;
};

// A leaf of the adaptive quadtree: level, box number (children 4b + c, as
// the uniform tree numbers them) and centre.
struct AdaptiveLeaf {
  int level;
  int64_t num;
  double cx, cy;
};
// The leaves in the reference's leafNodes order; *n_levels = the tree depth.
std::vector<AdaptiveLeaf> adaptive_leaves(const Params& P, int tree_adaptivity, int* n_levels);
// getIfNeighbor between target leaf a (radius ra) and source leaf b (rb):
// -1 far, else the family (0 colleague, 1 fine, 2 separated fine, 3 coarse)
// with its slot.
int classify_near(const AdaptiveLeaf& a, const AdaptiveLeaf& b, double ra, double rb, int& slot);

// One precomputation table: family -1 the far M at level (NTERM x R);
// 0..3 the colleague / fine / separated-fine / coarse neighbour operator n of
// a target box at level (R x R) (LS_Assembly.hpp:1328-1367).
struct TableSpec {
  int family, level, n;
};
extern const int FAMILY_SIZE[4];
constexpr int FAR_LEVELS = 10;  // M is precomputed for levels 0..10 (FarFieldInteraction.hpp:38)

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
  // precomp_dir empty: tables by device quadrature; else the tables are read
  // from the precomputation cache under precomp_dir (the reference's
  // precomputations/ text files; LS_Assembly.hpp:1217-1228, 1422-1459)
  explicit Problem(const Params& P, const std::string& precomp_dir = std::string());
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
  // adaptive discretisation: its leaves (nleaf of them, L = the depth), and
  // per leaf the near leaves with their table (family, level, slot) index
  std::vector<AdaptiveLeaf> leaves;
  std::vector<int> near_off, near_leaf, near_tab;
  std::vector<TableSpec> near_specs;
  void need_uniform(const char* what) const;
  double t_tables = 0, t_nca = 0, t_assemble = 0;
  Arena persist{size_t(64) << 20};
  Arena op{size_t(1) << 30};
  Arena scratch{size_t(256) << 20};
  Arena acaws{size_t(1) << 30};
  Arena acakeep{size_t(256) << 20};
  hipStream_t stream = nullptr;

  void build_tables();
  void load_tables(const std::string& dir);
  void install_tables();
  double box_radius(int j) const;
  std::vector<std::vector<double2>> quad_tables(const std::vector<TableSpec>& specs, int* capped);
  // every table of the reference's precomputation (M at levels 0..10, the 53
  // neighbour operators at levels 2..L) in the text layout under dir;
  // returns the seconds spent on quadrature
  double write_precomputations(const std::string& dir);
  double t_precompute = 0;
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
  void init(const std::string& precomp_dir);
  void init_adaptive();
  void nca_level(int j);
};

}  // namespace aifmm
