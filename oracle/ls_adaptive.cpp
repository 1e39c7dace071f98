// ORACLE — test infrastructure only (see ls_oracle.hpp).
// The adaptive quadtree with 2:1 level restriction and getMatrixEntry between
// its leaves, restated from LS/LS_Assembly.hpp (the reference builds it in
// LS2DTree's ctor when createAdaptiveTree + makeLevelRestriction_Tree are
// called instead of createUniformTree, :255-257).  The box lists, index
// searches and angle tests follow the reference's own structure.
#include <algorithm>
#include <cmath>
#include <stdexcept>
#include <string>

#include "ls_oracle.hpp"

namespace orc {

static const double PI = 3.1415926535897932384;  // LS_Assembly.hpp:30
static const double epsilonRoundOff = 1e-16;     // LS_Assembly.hpp:26

// LS_Assembly.hpp:73-120
static const double SFN_centers[20][2] = {{-2.5, -2.5}, {-1.5, -2.5}, {-0.5, -2.5}, {0.5, -2.5}, {1.5, -2.5},
                                          {2.5, -2.5},  {2.5, -1.5},  {2.5, -0.5},  {2.5, 0.5},  {2.5, 1.5},
                                          {2.5, 2.5},   {1.5, 2.5},   {0.5, 2.5},   {-0.5, 2.5}, {-1.5, 2.5},
                                          {-2.5, 2.5},  {-2.5, 1.5},  {-2.5, 0.5},  {-2.5, -0.5}, {-2.5, -1.5}};
static const double N_centers[9][2] = {{-2, -2}, {0, -2}, {2, -2}, {2, 0}, {2, 2}, {0, 2}, {-2, 2}, {-2, 0}, {0, 0}};
static const double FN_centers[12][2] = {{-1.5, -1.5}, {-0.5, -1.5}, {0.5, -1.5}, {1.5, -1.5},
                                         {1.5, -0.5},  {1.5, 0.5},   {1.5, 1.5},  {0.5, 1.5},
                                         {-0.5, 1.5},  {-1.5, 1.5},  {-1.5, 0.5}, {-1.5, -0.5}};
static const double CN_centers[12][2] = {{-3, -3}, {-1, -3}, {1, -3}, {3, -3}, {3, -1}, {3, 1},
                                         {3, 3},   {1, 3},   {-1, 3}, {-3, 3}, {-3, 1}, {-3, -1}};

namespace {

double get_S(double x, double y, int n) {  // LS_Assembly.hpp:553-565
  double S = 0.5;
  for (int k = 1; k < n; ++k) S += cos(k * acos(x)) * cos(k * acos(y));
  return 2.0 / n * S;
}

struct Builder {
  AdaptiveDisc& A;
  const Params& P;
  int rank, nc;
  std::vector<double> sx, sy, cx_, cy_;  // standardChebNodes, standardChebNodesChild
  std::vector<Mat> L2L;                  // real, stored in the real parts

  Builder(AdaptiveDisc& a, const Params& p) : A(a), P(p) {
    nc = P.n_cheb;
    rank = nc * nc;
    // set_Standard_Cheb_Nodes (LS_Assembly.hpp:567-607)
    std::vector<double> c1(nc);
    for (int k = 0; k < nc; ++k) c1[k] = -cos((k + 0.5) / nc * PI);
    for (int j = 0; j < nc; ++j)
      for (int k = 0; k < nc; ++k) sx.push_back(c1[k]), sy.push_back(c1[j]);
    const double off[4][2] = {{-0.5, -0.5}, {0.5, -0.5}, {0.5, 0.5}, {-0.5, 0.5}};
    for (int c = 0; c < 4; ++c)
      for (int j = 0; j < rank; ++j) cx_.push_back(0.5 * sx[j] + off[c][0]), cy_.push_back(0.5 * sy[j] + off[c][1]);
    // get_Transfer_Matrix (LS_Assembly.hpp:609-626)
    for (int l = 0; l < 4; ++l) {
      Mat m(rank, rank);
      for (int j = 0; j < rank; ++j)
        for (int k = 0; k < rank; ++k) m(j, k) = get_S(sx[k], cx_[j + l * rank], nc) * get_S(sy[k], cy_[j + l * rank], nc);
      L2L.push_back(m);
    }
  }

  // resolveContrast (LS_Assembly.hpp:628-643) with errInApproximation (:1132-1140)
  double resolveContrast(int j, int k) {
    const AdaptiveDisc::ABox& b = A.tree[j][k];
    std::vector<double> qn(rank);
    for (int i = 0; i < rank; ++i) qn[i] = contrast(P, b.radius * sx[i] + b.cx, b.radius * sy[i] + b.cy);
    double err = 0;
    for (int c = 0; c < 4; ++c)
      for (int jj = 0; jj < rank; ++jj) {
        double ap = 0;
        for (int kk = 0; kk < rank; ++kk) ap += L2L[c](jj, kk).real() * qn[kk];
        double tv = contrast(P, b.radius * cx_[jj + c * rank] + b.cx, b.radius * cy_[jj + c * rank] + b.cy);
        err = std::max(err, fabs(tv - ap));
      }
    return err;
  }

  void push_children(int j, int k, bool leaf) {
    if ((int)A.tree.size() == j + 1) A.tree.emplace_back(), A.indexTree.emplace_back();
    const AdaptiveDisc::ABox p = A.tree[j][k];
    for (int c = 0; c < 4; ++c) {
      AdaptiveDisc::ABox box;
      box.isLeaf = leaf;
      box.level = j + 1;
      box.boxNumber = p.boxNumber * 4 + c;
      box.radius = 0.5 * p.radius;
      box.cx = (c == 0 || c == 3) ? p.cx - 0.5 * p.radius : p.cx + 0.5 * p.radius;
      box.cy = (c == 0 || c == 1) ? p.cy - 0.5 * p.radius : p.cy + 0.5 * p.radius;
      A.tree[j + 1].push_back(box);
      A.indexTree[j + 1].push_back(box.boxNumber);
    }
  }

  // createTree (LS_Assembly.hpp:663-719); condition2Leaf is initialised true
  // there, so the right-hand-side test never decides
  void createTree(int j, int k) {
    double errContrast = resolveContrast(j, k);
    bool condition1Leaf = errContrast < A.epsilon_tree;
    bool condition3Leaf = P.kappa * A.tree[j][k].radius / 2 / PI <= 1.0;
    if (condition1Leaf && condition3Leaf) {
      A.tree[j][k].isLeaf = true;
      A.leafNodes.push_back({j, k});
    } else {
      int n = (int)A.tree.size() == j + 1 ? 0 : (int)A.tree[j + 1].size();
      push_children(j, k, false);
      for (int c = 0; c < 4; ++c) createTree(j + 1, n + c);
    }
  }

  int index_of(int level, int boxNumber) const {  // std::find in indexTree
    auto it = std::find(A.indexTree[level].begin(), A.indexTree[level].end(), boxNumber);
    return (int)(it - A.indexTree[level].begin());
  }

  using NB = std::vector<std::vector<std::vector<std::pair<int, int>>>>;

  // getFutureGeneration_LFR (LS_Assembly.hpp:892-912)
  void future(int bj, int bk, int pnj, int pni, std::vector<std::pair<int, int>>& fg) {
    const auto& b = A.tree[bj][bk];
    const auto& p = A.tree[pnj][pni];
    if (p.isLeaf) {
      if (fabs(b.cx - p.cx) >= 3.0 * A.boxRadius[bj] + A.boxRadius[pnj] - epsilonRoundOff ||
          fabs(b.cy - p.cy) >= 3.0 * A.boxRadius[bj] + A.boxRadius[pnj] - epsilonRoundOff) {
      } else {
        fg.push_back({pnj, pni});
      }
    } else {
      for (int nc2 = 0; nc2 < 4; ++nc2) future(bj, bk, pnj + 1, index_of(pnj + 1, 4 * p.boxNumber + nc2), fg);
    }
  }

  // assign_Tree_Interactions_LFR (LS_Assembly.hpp:914-966)
  void interactions(NB& T) {
    for (int c = 0; c < 4; ++c)
      for (int n = 0; n < 4; ++n) T[1][c].push_back({1, n});
    for (int j = 1; j <= A.n_levels - 1; ++j)
      for (int k = 0; k < (int)A.tree[j].size(); ++k) {
        if (A.tree[j][k].isLeaf) continue;
        for (int c = 0; c < 4; ++c) {
          int boxA_index = index_of(j + 1, 4 * A.tree[j][k].boxNumber + c);
          for (size_t n = 0; n < T[j][k].size(); ++n) {
            std::vector<std::pair<int, int>> fg;
            future(j + 1, boxA_index, T[j][k][n].first, T[j][k][n].second, fg);
            for (auto& e : fg) T[j + 1][boxA_index].push_back(e);
          }
        }
      }
  }

  int getIndexOfInLeafNodes(std::pair<int, int> op) {  // LS_Assembly.hpp:1003-1009
    for (size_t l = 0; l < A.leafNodes.size(); ++l)
      if (A.leafNodes[l] == op) return (int)l;
    throw std::runtime_error("leaf not found");
  }

  // refineBox (LS_Assembly.hpp:1036-1072)
  void refineBox(int l, std::pair<int, int> leaf, std::vector<std::pair<int, int>>& old) {
    int j = leaf.first, k = leaf.second;
    ++A.restriction_splits;
    old[l].first = -1;
    A.tree[j][k].isLeaf = false;
    int n = (int)A.tree.size() == j + 1 ? 0 : (int)A.tree[j + 1].size();
    push_children(j, k, true);
    for (int c = 0; c < 4; ++c) A.leafNodes.push_back({j + 1, n + c});
  }

  // makeLevelRestriction_Tree / _Box (LS_Assembly.hpp:968-1034)
  void makeLevelRestriction() {
    for (int pass = 0; pass < A.n_levels; ++pass) {
      NB T(A.n_levels + 1);
      for (int j = 0; j <= A.n_levels; ++j) T[j].resize(A.tree[j].size());
      interactions(T);
      std::vector<std::pair<int, int>> old = A.leafNodes;
      for (size_t l = 0; l < old.size(); ++l) {
        std::pair<int, int> leaf = A.leafNodes[l];
        int j = leaf.first, k = leaf.second;
        for (size_t i = 0; i < T[j][k].size(); ++i) {
          int nj = T[j][k][i].first;
          if (j - nj >= 2) {
            int idx = getIndexOfInLeafNodes(T[j][k][i]);
            if (old[idx].first != -1) refineBox(idx, T[j][k][i], old);
          } else if (nj - j >= 2) {
            if (old[l].first != -1) refineBox((int)l, leaf, old);
          }
        }
      }
      int cnt = 0;
      for (size_t l = 0; l < old.size(); ++l)
        if (old[l].first == -1) {
          A.leafNodes.erase(A.leafNodes.begin() + l - cnt);
          cnt++;
        }
    }
  }
};

}  // namespace

void AdaptiveDisc::build(const Params& Pin, int tree_adaptivity) {
  D.init_basis(Pin);
  const Params& P = D.P;
  rank = D.rank;
  epsilon_tree = pow(10.0, -tree_adaptivity);  // LS_Assembly.hpp:201
  tree.clear();
  indexTree.clear();
  leafNodes.clear();
  Builder B(*this, P);
  // createAdaptiveTree (LS_Assembly.hpp:721-742)
  ABox root;
  root.level = 0;
  root.boxNumber = 0;
  root.radius = P.L;
  tree.push_back({root});
  indexTree.push_back({0});
  B.createTree(0, 0);
  n_levels = (int)tree.size() - 1;
  boxRadius.assign(1, P.L);
  for (int j = 1; j <= std::max(10, n_levels + 1); ++j) boxRadius.push_back(boxRadius[j - 1] / 2.0);
  B.makeLevelRestriction();
  // assignLeafChargeLocations (LS_Assembly.hpp:1074-1087)
  N = (int64_t)leafNodes.size() * rank;
  px.clear();
  py.clear();
  for (auto& lf : leafNodes) {
    const ABox& b = tree[lf.first][lf.second];
    for (int i = 0; i < rank; ++i) {
      px.push_back(boxRadius[lf.first] * B.sx[i] + b.cx);
      py.push_back(boxRadius[lf.first] * B.sy[i] + b.cy);
    }
  }
  q.resize(N);
  for (int64_t i = 0; i < N; ++i) q[i] = contrast(P, px[i], py[i]);
}

// getIfNeighbor (LS_Assembly.hpp:283-472) for the leaves' boxes, with the
// angle tables of getNeighbors (:1269-1288)
int AdaptiveDisc::neighbor(int li, int lj, int& num) const {
  const ABox& a = tree[leafNodes[li].first][leafNodes[li].second];
  const ABox& b = tree[leafNodes[lj].first][leafNodes[lj].second];
  num = -1;
  const int j = a.level, nj = b.level;
  if (fabs(a.cx - b.cx) >= 3.0 * boxRadius[j] + boxRadius[nj] - epsilonRoundOff ||
      fabs(a.cy - b.cy) >= 3.0 * boxRadius[j] + boxRadius[nj] - epsilonRoundOff)
    return -1;
  const double angle = atan2(b.cy - a.cy, b.cx - a.cx);
  auto match = [&](const double (*c)[2], int n) {
    for (int s = 0; s < n; ++s)
      if (fabs(atan2(c[s][1], c[s][0]) - angle) < epsilonRoundOff) return s;
    return -1;
  };
  if (a.radius == b.radius) {
    if (fabs(b.cy - a.cy) < epsilonRoundOff && fabs(b.cx - a.cx) < epsilonRoundOff) num = 8;
    else num = match(N_centers, 8);
    return 0;
  }
  if (a.radius < b.radius) {
    num = match(CN_centers, 12);
    return 1;
  }
  const double dist = sqrt((a.cx - b.cx) * (a.cx - b.cx) + (a.cy - b.cy) * (a.cy - b.cy));  // find_distance
  if (dist > 2.5 * a.radius) {
    num = match(SFN_centers, 19);
    if (num < 0) num = 19;
    return 3;
  }
  num = match(FN_centers, 12);
  return 2;
}

// The tables of LS_Assembly.hpp:1328-1367 (source level = target level + 1
// for fine and separated-fine, - 1 for coarse), computed once each and
// rounded as the text cache would.  Disc::table's families: 0 colleague, 1
// fine, 2 separated fine, 3 coarse.
const Mat& AdaptiveDisc::near_table(int type, int level, int num) {
  static const int fam_of_type[4] = {0, 3, 1, 2};
  std::vector<int> key{type, level, num};
  auto it = ntab_.find(key);
  if (it != ntab_.end()) return it->second;
  Mat t = D.table(fam_of_type[type], level, num);
  if (D.P.quantize)
    for (auto& v : t.a) v = round6c(v);
  return ntab_[key] = t;
}

const Mat& AdaptiveDisc::far_table(int level) {
  auto it = ftab_.find(level);
  if (it != ftab_.end()) return it->second;
  Mat t = D.table(-1, level, 0);
  if (D.P.quantize)
    for (auto& v : t.a) v = round6c(v);
  return ftab_[level] = t;
}

// getMatrixEntry (LS_Assembly.hpp:474-541)
cd AdaptiveDisc::entry(int64_t i, int64_t j) {
  const int leaf_i = (int)(i / rank), index_i = (int)(i % rank);
  const int leaf_j = (int)(j / rank), index_j = (int)(j % rank);
  const int level_i = leafNodes[leaf_i].first;
  const ABox& bj = tree[leafNodes[leaf_j].first][leafNodes[leaf_j].second];
  int num;
  const int type = neighbor(leaf_i, leaf_j, num);
  const double kappa = D.P.kappa;
  if (type != -1) {
    if (num < 0) throw std::runtime_error("adaptive oracle: near pair matches no neighbour slot");
    cd e = near_table(type, level_i, num)(index_i, index_j);
    if (find_phi) {
      e = e * kappa * kappa * q[i];
      if (i == j) e = 1.0 + e;
    }
    if (std::abs(e) < 1e-16) e = 0.0;
    return e;
  }
  cd s = graf_series(far_table(bj.level), index_j, px[i], py[i], bj.cx, bj.cy, kappa, D.series,
                     D.P.exact_bessel != 0);
  s *= cd(0.0, 0.25);
  if (find_phi) {
    s *= kappa * kappa * q[i];
    if (i == j) s += 1.0;
  }
  if (std::abs(s) < 1e-16) s = 0.0;
  return s;
}

}  // namespace orc
