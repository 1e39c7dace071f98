// Adaptive quadtree with 2:1 level restriction: LS2DTree::createAdaptiveTree
// and makeLevelRestriction_Tree (LS_Assembly.hpp:609-742, 892-1072), and the
// neighbour typing of getMatrixEntry between leaves of different levels
// (getIfNeighbor, LS_Assembly.hpp:283-472).  Host code: the tree is built once
// per problem, then the device evaluates entries and products over its leaves.
#include <cmath>
#include <stdexcept>
#include <string>
#include <unordered_map>

#include "problem.h"

namespace aifmm {

static const double PI = 3.1415926535897932384;  // LS_Assembly.hpp:30
static const double kRoundOff = 1e-16;           // epsilonRoundOff, LS_Assembly.hpp:26

namespace {

// Chebyshev interpolation weight between nodes of n points
// (get_S, LS_Assembly.hpp:553-565)
double cheb_s(double x, double y, int n) {
  double s = 0.5;
  for (int k = 1; k < n; ++k) s += cos(k * acos(x)) * cos(k * acos(y));
  return 2.0 / n * s;
}

struct TBox {
  int64_t num;
  double cx, cy;
  bool leaf = false;
};

struct Builder {
  const Params& P;
  double eps_tree;
  int nc, rank;
  std::vector<double> xi;                 // standard nodes, 2D: index jy*nc + kx
  std::vector<double> cxn, cyn;           // child nodes in the parent's frame, 4 x rank
  std::vector<std::vector<double>> l2l;   // 4 interpolation matrices, rank x rank (row-major [j][k])
  std::vector<std::vector<TBox>> tree;    // per level, in creation order
  std::vector<std::unordered_map<int64_t, int>> at;  // box number -> index, per level
  std::vector<std::pair<int, int>> leaves;  // (level, index): the reference's leafNodes

  Builder(const Params& p, int tree_adaptivity) : P(p) {
    eps_tree = pow(10.0, -tree_adaptivity);
    nc = P.n_cheb;
    rank = nc * nc;
    std::vector<double> c1(nc);
    for (int k = 0; k < nc; ++k) c1[k] = -cos((k + 0.5) / nc * PI);
    // set_Standard_Cheb_Nodes (LS_Assembly.hpp:567-607): x fastest; children
    // 0 (-,-), 1 (+,-), 2 (+,+), 3 (-,+)
    std::vector<double> sx(rank), sy(rank);
    for (int j = 0; j < nc; ++j)
      for (int k = 0; k < nc; ++k) sx[j * nc + k] = c1[k], sy[j * nc + k] = c1[j];
    static const double cs[4][2] = {{-0.5, -0.5}, {0.5, -0.5}, {0.5, 0.5}, {-0.5, 0.5}};
    cxn.resize(4 * rank);
    cyn.resize(4 * rank);
    for (int c = 0; c < 4; ++c)
      for (int j = 0; j < rank; ++j) {
        cxn[c * rank + j] = 0.5 * sx[j] + cs[c][0];
        cyn[c * rank + j] = 0.5 * sy[j] + cs[c][1];
      }
    // get_Transfer_Matrix (LS_Assembly.hpp:609-626)
    l2l.assign(4, std::vector<double>((size_t)rank * rank));
    for (int c = 0; c < 4; ++c)
      for (int j = 0; j < rank; ++j)
        for (int k = 0; k < rank; ++k)
          l2l[c][(size_t)j * rank + k] = cheb_s(sx[k], cxn[c * rank + j], nc) * cheb_s(sy[k], cyn[c * rank + j], nc);
    xi.resize(2 * rank);
    for (int j = 0; j < rank; ++j) xi[2 * j] = sx[j], xi[2 * j + 1] = sy[j];
  }

  double radius(int j) const {  // tree[j][k].radius: the root's L halved j times
    double r = P.L;
    for (int i = 0; i < j; ++i) r = 0.5 * r;
    return r;
  }

  // resolveContrast (LS_Assembly.hpp:628-643): the largest error of
  // interpolating the contrast from the box's nodes to its children's nodes
  double contrast_error(int j, int k) const {
    const TBox& b = tree[j][k];
    const double r = radius(j);
    std::vector<double> qp(rank);
    for (int i = 0; i < rank; ++i) qp[i] = contrast(P, r * xi[2 * i] + b.cx, r * xi[2 * i + 1] + b.cy);
    double err = 0;
    for (int c = 0; c < 4; ++c)
      for (int jj = 0; jj < rank; ++jj) {
        double a = 0;
        for (int kk = 0; kk < rank; ++kk) a += l2l[c][(size_t)jj * rank + kk] * qp[kk];
        const double t = contrast(P, r * cxn[c * rank + jj] + b.cx, r * cyn[c * rank + jj] + b.cy);
        err = std::max(err, fabs(t - a));
      }
    return err;
  }

  void add_children(int j, int k) {
    if ((int)tree.size() == j + 1) tree.emplace_back(), at.emplace_back();
    const TBox b = tree[j][k];
    const double r = radius(j);
    static const double cs[4][2] = {{-0.5, -0.5}, {0.5, -0.5}, {0.5, 0.5}, {-0.5, 0.5}};
    for (int c = 0; c < 4; ++c) {
      TBox ch;
      ch.num = 4 * b.num + c;
      ch.cx = b.cx + cs[c][0] * r;
      ch.cy = b.cy + cs[c][1] * r;
      at[j + 1][ch.num] = (int)tree[j + 1].size();
      tree[j + 1].push_back(ch);
    }
  }

  // createTree (LS_Assembly.hpp:663-719): a box is a leaf when its contrast
  // is resolved and kappa r / (2 pi) <= 1 (the right-hand-side test is
  // always true there); otherwise its four children are created, then each
  // visited in turn
  void create(int j, int k) {
    const double err = contrast_error(j, k);
    const bool c1 = err < eps_tree;
    const bool c3 = P.kappa * radius(j) / 2 / PI <= 1.0;
    if (c1 && c3) {
      tree[j][k].leaf = true;
      leaves.push_back({j, k});
      return;
    }
    const int n = (int)(tree.size() > (size_t)j + 1 ? tree[j + 1].size() : 0);
    add_children(j, k);
    for (int c = 0; c < 4; ++c) create(j + 1, n + c);
  }

  // the near test of getFutureGeneration_LFR and getIfNeighbor
  bool near(int j, int k, int nj, int nk) const {
    const double lim = 3.0 * radius(j) + radius(nj) - kRoundOff;
    return !(fabs(tree[j][k].cx - tree[nj][nk].cx) >= lim || fabs(tree[j][k].cy - tree[nj][nk].cy) >= lim);
  }

  // getFutureGeneration_LFR (LS_Assembly.hpp:892-912): the leaves under
  // (pj, pk) near box (bj, bk)
  void future(int bj, int bk, int pj, int pk, std::vector<std::pair<int, int>>& out) const {
    if (tree[pj][pk].leaf) {
      if (near(bj, bk, pj, pk)) out.push_back({pj, pk});
      return;
    }
    for (int c = 0; c < 4; ++c) future(bj, bk, pj + 1, at[pj + 1].at(4 * tree[pj][pk].num + c), out);
  }

  // makeLevelRestriction_Tree (LS_Assembly.hpp:968-1001): n_levels passes;
  // each pass lists every leaf's near leaves (assign_Tree_Interactions_LFR),
  // then refines a leaf that has a near leaf two or more levels finer and
  // every near leaf two or more levels coarser; refined leaves leave the leaf
  // list, their children join its end
  void restrict_levels() {
    const int nlev = (int)tree.size() - 1;
    for (int pass = 0; pass < nlev; ++pass) {
      std::vector<std::vector<std::vector<std::pair<int, int>>>> nb(nlev + 1);
      for (int j = 0; j <= nlev; ++j) nb[j].resize(tree[j].size());
      if (nlev >= 1)
        for (int c = 0; c < 4; ++c)
          for (int n = 0; n < 4; ++n) nb[1][c].push_back({1, n});
      for (int j = 1; j <= nlev - 1; ++j)
        for (int k = 0; k < (int)tree[j].size(); ++k) {
          if (tree[j][k].leaf) continue;
          for (int c = 0; c < 4; ++c) {
            const int a = at[j + 1].at(4 * tree[j][k].num + c);
            for (auto& pn : nb[j][k]) future(j + 1, a, pn.first, pn.second, nb[j + 1][a]);
          }
        }
      std::vector<std::pair<int, int>> old = leaves;
      std::vector<char> gone(old.size(), 0);
      std::unordered_map<int64_t, int> pos;  // (level, index) -> position in old
      for (size_t l = 0; l < old.size(); ++l) pos[((int64_t)old[l].first << 40) | old[l].second] = (int)l;
      auto refine = [&](int l) {
        gone[l] = 1;
        const int j = old[l].first, k = old[l].second;
        tree[j][k].leaf = false;
        const int n = (int)(tree.size() > (size_t)j + 1 ? tree[j + 1].size() : 0);
        add_children(j, k);
        for (int c = 0; c < 4; ++c) {
          tree[j + 1][n + c].leaf = true;
          leaves.push_back({j + 1, n + c});
        }
      };
      for (size_t l = 0; l < old.size(); ++l) {
        const int j = old[l].first, k = old[l].second;
        for (auto& q : nb[j][k]) {
          if (j - q.first >= 2) {
            const int ql = pos.at(((int64_t)q.first << 40) | q.second);
            if (!gone[ql]) refine(ql);
          } else if (q.first - j >= 2) {
            if (!gone[l]) refine((int)l);
          }
        }
      }
      std::vector<std::pair<int, int>> kept;
      for (size_t l = 0; l < leaves.size(); ++l)
        if (l >= old.size() || !gone[l]) kept.push_back(leaves[l]);
      leaves.swap(kept);
    }
  }
};

}  // namespace

std::vector<AdaptiveLeaf> adaptive_leaves(const Params& P, int tree_adaptivity, int* n_levels) {
  if (tree_adaptivity < 0 || tree_adaptivity > 15) throw std::runtime_error("tree adaptivity must be in [0, 15]");
  Builder b(P, tree_adaptivity);
  b.tree.emplace_back();
  b.at.emplace_back();
  TBox root;
  root.num = 0;
  root.cx = root.cy = 0.0;
  b.tree[0].push_back(root);
  b.at[0][0] = 0;
  b.create(0, 0);
  if ((int)b.tree.size() - 1 > 12) throw std::runtime_error("adaptive tree deeper than 12 levels");
  b.restrict_levels();
  if (n_levels) *n_levels = (int)b.tree.size() - 1;
  std::vector<AdaptiveLeaf> out;
  out.reserve(b.leaves.size());
  for (auto& l : b.leaves) {
    const TBox& t = b.tree[l.first][l.second];
    out.push_back(AdaptiveLeaf{l.first, t.num, t.cx, t.cy});
  }
  return out;
}

// Source-box centres of the neighbour families in units of the target box
// radius (LS_Assembly.hpp:73-120), in the reference's slot order.
static const double NC[9][2] = {{-2, -2}, {0, -2}, {2, -2}, {2, 0}, {2, 2}, {0, 2}, {-2, 2}, {-2, 0}, {0, 0}};
static const double FNC[12][2] = {{-1.5, -1.5}, {-0.5, -1.5}, {0.5, -1.5}, {1.5, -1.5}, {1.5, -0.5}, {1.5, 0.5},
                                  {1.5, 1.5},   {0.5, 1.5},   {-0.5, 1.5}, {-1.5, 1.5}, {-1.5, 0.5}, {-1.5, -0.5}};
static const double SFNC[20][2] = {{-2.5, -2.5}, {-1.5, -2.5}, {-0.5, -2.5}, {0.5, -2.5}, {1.5, -2.5},
                                   {2.5, -2.5},  {2.5, -1.5},  {2.5, -0.5},  {2.5, 0.5},  {2.5, 1.5},
                                   {2.5, 2.5},   {1.5, 2.5},   {0.5, 2.5},   {-0.5, 2.5}, {-1.5, 2.5},
                                   {-2.5, 2.5},  {-2.5, 1.5},  {-2.5, 0.5},  {-2.5, -0.5}, {-2.5, -1.5}};
static const double CNC[12][2] = {{-3, -3}, {-1, -3}, {1, -3}, {3, -3}, {3, -1}, {3, 1},
                                  {3, 3},   {1, 3},   {-1, 3}, {-3, 3}, {-3, 1}, {-3, -1}};

// getIfNeighbor (LS_Assembly.hpp:283-472) for leaves a (target) and b
// (source): -1 far; else the family (0 colleague, 1 fine, 2 separated fine,
// 3 coarse; the reference's typeOfNeighbor 0, 2, 3, 1) and its slot.  The
// reference matches atan2 of the centre offset against the families'
// centres to 1e-16; here the offset in units of the target radius is matched
// to the centres, which names the same slot for every pair of a tree whose
// offsets are exact multiples of the radii (L a power of two) and stays
// defined otherwise.  An unmatched separated-fine pair takes slot 19, as the
// reference's final else does; any other unmatched pair is an error.
int classify_near(const AdaptiveLeaf& a, const AdaptiveLeaf& b, double ra, double rb, int& slot) {
  const double dx = b.cx - a.cx, dy = b.cy - a.cy;
  const double lim = 3.0 * ra + rb - kRoundOff;
  if (fabs(dx) >= lim || fabs(dy) >= lim) return -1;
  const double ox = dx / ra, oy = dy / ra;
  auto match = [&](const double (*c)[2], int n) {
    for (int s = 0; s < n; ++s)
      if (fabs(c[s][0] - ox) < 1e-9 && fabs(c[s][1] - oy) < 1e-9) return s;
    return -1;
  };
  int fam;
  if (a.level == b.level) {
    fam = 0;
    slot = match(NC, 9);
  } else if (b.level < a.level) {
    fam = 3;
    slot = match(CNC, 12);
  } else if (sqrt(dx * dx + dy * dy) > 2.5 * ra) {
    fam = 2;
    slot = match(SFNC, 20);
    if (slot < 0) slot = 19;
  } else {
    fam = 1;
    slot = match(FNC, 12);
  }
  if (slot < 0)
    throw std::runtime_error("adaptive tree: near leaves at levels " + std::to_string(a.level) + " and " +
                             std::to_string(b.level) + " match no neighbour slot");
  return fam;
}

}  // namespace aifmm
