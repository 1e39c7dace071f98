// ORACLE — test infrastructure only (see ls_oracle.hpp).
// Restatement of FMM2DTree<userkernel> (FMM2DTree_find_error.hpp) and
// LowRank::ACA_only_nodes (ACA.hpp): quadtree lists, NCA skeletons, M2L and
// P2P assembly, the elimination schedules, the base solve and the
// back-substitution.
#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <string>

#include "ls_oracle.hpp"

namespace orc {

// neighborNumbers slots 0..7 = SW,S,SE,E,NE,N,NW,W
// (FMM2DTree_find_error.hpp:168-1046, drawings in assign_Child*_Interaction)
const int NB_OFF[8][2] = {{-1, -1}, {0, -1}, {1, -1}, {1, 0}, {1, 1}, {0, 1}, {-1, 1}, {-1, 0}};
// innerNumbers: distance-2 ring, counter-clockwise from (-2,-2)
const int INNER_OFF[16][2] = {{-2, -2}, {-1, -2}, {0, -2}, {1, -2}, {2, -2}, {2, -1}, {2, 0}, {2, 1},
                              {2, 2},   {1, 2},   {0, 2},  {-1, 2}, {-2, 2}, {-2, 1}, {-2, 0}, {-2, -1}};
// outerNumbers: distance-3 ring, counter-clockwise from (-3,-3)
const int OUTER_OFF[24][2] = {{-3, -3}, {-2, -3}, {-1, -3}, {0, -3}, {1, -3}, {2, -3}, {3, -3}, {3, -2},
                              {3, -1},  {3, 0},   {3, 1},   {3, 2},  {3, 3},  {2, 3},  {1, 3},  {0, 3},
                              {-1, 3},  {-2, 3},  {-3, 3},  {-3, 2}, {-3, 1}, {-3, 0}, {-3, -1}, {-3, -2}};

void box_xy(int level, int b, int& ix, int& iy) {
  ix = iy = 0;
  for (int t = level - 1; t >= 0; --t) {
    int c = (b >> (2 * t)) & 3;
    ix = 2 * ix + ((c == 1 || c == 2) ? 1 : 0);
    iy = 2 * iy + ((c == 2 || c == 3) ? 1 : 0);
  }
}

int xy_box(int level, int ix, int iy) {
  int b = 0;
  for (int t = level - 1; t >= 0; --t) {
    int bx = (ix >> t) & 1, by = (iy >> t) & 1;
    int c = by ? (bx ? 2 : 3) : (bx ? 1 : 0);
    b = 4 * b + c;
  }
  return b;
}

void Tree::create(const Disc* Din, int tp) {
  D = Din;
  L = D->P.n_levels;
  tol_pow = tp;
  rrqr_thr = pow(10, -1.0 * tp);  // FMM2DTree_find_error.hpp:115
  T.assign(L + 1, {});
  order.assign(L + 1, {});
  for (int j = 0; j <= L; ++j) {
    int nb = 1 << (2 * j), side = 1 << j;
    T[j].resize(nb);
    for (int k = 0; k < nb; ++k) {
      Box& B = T[j][k];
      for (int s = 0; s < 8; ++s) B.nb[s] = -1;
      for (int s = 0; s < 16; ++s) B.inner[s] = -1;
      for (int s = 0; s < 24; ++s) B.outer[s] = -1;
      if (j == 0) continue;
      int ix, iy;
      box_xy(j, k, ix, iy);
      auto inside = [&](int x, int y) { return x >= 0 && y >= 0 && x < side && y < side; };
      for (int s = 0; s < 8; ++s) {
        int x = ix + NB_OFF[s][0], y = iy + NB_OFF[s][1];
        if (inside(x, y)) B.nb[s] = xy_box(j, x, y);
      }
      for (int s = 0; s < 16; ++s) {
        int x = ix + INNER_OFF[s][0], y = iy + INNER_OFF[s][1];
        if (inside(x, y)) B.inner[s] = xy_box(j, x, y);
      }
      for (int s = 0; s < 24; ++s) {
        int x = ix + OUTER_OFF[s][0], y = iy + OUTER_OFF[s][1];
        if (inside(x, y) && std::abs((x >> 1) - (ix >> 1)) <= 1 && std::abs((y >> 1) - (iy >> 1)) <= 1)
          B.outer[s] = xy_box(j, x, y);
      }
    }
  }
  // leaf charges (FMM2DTree_find_error.hpp:1072-1090)
  int r = D->rank;
  for (int k = 0; k < (1 << (2 * L)); ++k)
    for (int i = 0; i < r; ++i) T[L][k].charges.push_back(k * r + i);
}

void Tree::assign_rhs(const cd* b) {  // FMM2DTree_find_error.hpp:1092-1103
  int r = D->rank;
  for (int k = 0; k < (1 << (2 * L)); ++k) {
    Box& B = T[L][k];
    B.particle_rhs = Mat(r, 1);
    for (int i = 0; i < r; ++i) B.particle_rhs(i, 0) = b[(int64_t)k * r + i];
  }
}

// ------------------------------------------------------------------ ACA
// LowRank::ACA_only_nodes (ACA.hpp:368-577) with maxAbsVector (ACA.hpp:27-40).
// maxAbsVector returns the first allowed index attaining the maximum modulus.
// Mirror-symmetric geometry makes exact ties common, and which of two tied
// entries wins then depends on last-bit rounding; ties are therefore taken
// within a relative 1e-12 of the maximum, still to the lowest index, so that
// every FP64 implementation picks the same pivot.
static const double ACA_TIE = 1e-12;
static void maxabs(const std::vector<cd>& v, const std::vector<char>& allowed, int& index) {
  double mx = -1;
  for (size_t i = 0; i < allowed.size(); ++i)
    if (allowed[i] && std::abs(v[i]) > mx) mx = std::abs(v[i]);
  if (mx < 0) {  // empty set: libstdc++ header read yields 0
    index = 0;
    return;
  }
  double thr = mx * (1.0 - ACA_TIE);
  for (size_t i = 0; i < allowed.size(); ++i)
    if (allowed[i] && std::abs(v[i]) >= thr) {
      index = (int)i;
      return;
    }
  index = 0;
}

static double vnorm(const std::vector<cd>& v) {
  double s = 0;
  for (auto& x : v) s += std::norm(x);
  return std::sqrt(s);
}
static cd vdot(const std::vector<cd>& a, const std::vector<cd>& b) {
  cd s = 0;
  for (size_t i = 0; i < a.size(); ++i) s += std::conj(a[i]) * b[i];
  return s;
}

void Tree::aca(const std::vector<int>& rows, const std::vector<int>& cols, std::vector<int>& rb,
               std::vector<int>& cb, Mat& Ac, Mat& Ar) const {
  const double tol = pow(10, -1.0 * tol_pow);
  const double tol_robust = pow(10.0, -12.0);
  int N1 = rows.size(), N2 = cols.size();
  auto getRow = [&](int ri) {
    std::vector<cd> v(N2);
    for (int c = 0; c < N2; ++c) v[c] = D->entry(rows[ri], cols[c]);
    return v;
  };
  auto getCol = [&](int ci) {
    std::vector<cd> v(N1);
    for (int r = 0; r < N1; ++r) v[r] = D->entry(rows[r], cols[ci]);
    return v;
  };
  rb.clear();
  cb.clear();
  std::vector<std::vector<cd>> U, V, AcV, ArV;
  int rank = 0;
  int mn = std::min(N1, N2);
  std::vector<char> remR(N1, 1), remC(N2, 1);
  int row_index = 0, col_index = 0;
  std::vector<cd> row, col, u, v;
  if (N1 < N2) {
    int l;
    for (l = 0; l < N1; ++l) {
      row_index = l;
      row = getRow(row_index);
      maxabs(row, remC, col_index);
      if (std::abs(row[col_index]) > tol_robust) break;
    }
    if (l == N1) {
      Ac = Mat(N1, 0);
      Ar = Mat(0, N2);
      return;
    }
    v.resize(N2);
    cd piv = row[col_index];
    for (int s = 0; s < N2; ++s) v[s] = row[s] / piv;
    rb.push_back(row_index);
    cb.push_back(col_index);
    col = getCol(col_index);
    u = col;
    U.push_back(u);
    V.push_back(v);
    AcV.push_back(col);
    ArV.push_back(row);
    remC[col_index] = 0;
    remR[row_index] = 0;
    rank = 1;
    double normS = 0.0, prev = DBL_MAX;
    maxabs(col, remR, row_index);
    while (std::abs(row[col_index]) > tol && std::abs(col[row_index]) > tol &&
           std::abs(prev - normS) >= tol * prev && rank < mn) {
      rb.push_back(row_index);
      row = getRow(row_index);
      std::vector<cd> row_temp = row;
      for (int q = 0; q < rank; ++q)
        for (int s = 0; s < N2; ++s) row[s] -= U[q][row_index] * V[q][s];
      maxabs(row, remC, col_index);
      cb.push_back(col_index);
      cd pv = row[col_index];
      for (int s = 0; s < N2; ++s) v[s] = row[s] / pv;
      col = getCol(col_index);
      std::vector<cd> col_temp = col;
      for (int q = 0; q < rank; ++q)
        for (int s = 0; s < N1; ++s) col[s] -= V[q][col_index] * U[q][s];
      u = col;
      if (vnorm(u) < tol_robust || vnorm(v) < tol_robust) {
        rb.pop_back();
        cb.pop_back();
        break;
      }
      U.push_back(u);
      V.push_back(v);
      AcV.push_back(col_temp);
      ArV.push_back(row_temp);
      ++rank;
      remC[col_index] = 0;
      remR[row_index] = 0;
      if (rank != 2) prev = normS;
      normS = pow(normS, 2) + pow(vnorm(u) * vnorm(v), 2);
      for (int q = 0; q < rank; ++q) normS += 2 * std::abs(vdot(U[q], u) * vdot(V[q], v));
      normS = sqrt(normS);
      maxabs(col, remR, row_index);
    }
  } else {
    int l;
    for (l = 0; l < N2; ++l) {
      col_index = l;
      col = getCol(col_index);
      maxabs(col, remR, row_index);
      if (std::abs(col[row_index]) > tol_robust) break;
    }
    if (l == N2) {
      Ac = Mat(N1, 0);
      Ar = Mat(0, N2);
      return;
    }
    cb.push_back(col_index);
    v.resize(N1);
    cd piv = col[row_index];
    for (int s = 0; s < N1; ++s) v[s] = col[s] / piv;
    rb.push_back(row_index);
    row = getRow(row_index);
    u = row;
    U.push_back(u);
    V.push_back(v);
    AcV.push_back(col);
    ArV.push_back(row);
    rank = 1;
    remR[row_index] = 0;
    remC[col_index] = 0;
    double normS = 0.0, prev = DBL_MAX;
    maxabs(row, remC, col_index);
    while (std::abs(row[col_index]) > tol && std::abs(col[row_index]) > tol &&
           std::abs(prev - normS) >= tol * prev && rank < mn) {
      cb.push_back(col_index);
      col = getCol(col_index);
      std::vector<cd> col_temp = col;
      for (int q = 0; q < rank; ++q)
        for (int s = 0; s < N1; ++s) col[s] -= U[q][col_index] * V[q][s];
      maxabs(col, remR, row_index);
      rb.push_back(row_index);
      cd pv = col[row_index];
      for (int s = 0; s < N1; ++s) v[s] = col[s] / pv;
      row = getRow(row_index);
      std::vector<cd> row_temp = row;
      for (int q = 0; q < rank; ++q)
        for (int s = 0; s < N2; ++s) row[s] -= V[q][row_index] * U[q][s];
      u = row;
      if (vnorm(u) < tol_robust || vnorm(v) < tol_robust) {
        cb.pop_back();
        rb.pop_back();
        break;
      }
      U.push_back(u);
      V.push_back(v);
      AcV.push_back(col_temp);
      ArV.push_back(row_temp);
      ++rank;
      remR[row_index] = 0;
      remC[col_index] = 0;
      if (rank != 2) prev = normS;
      normS = pow(normS, 2) + pow(vnorm(u) * vnorm(v), 2);
      for (int q = 0; q < rank; ++q) normS += 2 * std::abs(vdot(U[q], u) * vdot(V[q], v));
      normS = sqrt(normS);
      maxabs(row, remC, col_index);
    }
  }
  Ac = Mat(N1, rank);
  Ar = Mat(rank, N2);
  for (int q = 0; q < rank; ++q) {
    for (int i = 0; i < N1; ++i) Ac(i, q) = AcV[q][i];
    for (int i = 0; i < N2; ++i) Ar(q, i) = ArV[q][i];
  }
}

// ------------------------------------------------------------------ NCA
// getNodes_outgoing_box (FMM2DTree_find_error.hpp:1230-1281)
void Tree::nodes_outgoing(int j, int k) {
  Box& B = T[j][k];
  std::vector<int> boxA, IL;
  auto own_cols = [&](int kk, std::vector<int>& out) {
    if (j == L) out.insert(out.end(), T[j][kk].charges.begin(), T[j][kk].charges.end());
    else
      for (int c = 0; c < 4; ++c) {
        auto& ch = T[j + 1][4 * kk + c].out_chg;
        out.insert(out.end(), ch.begin(), ch.end());
      }
  };
  auto il_rows = [&](int kk, std::vector<int>& out) {
    if (j == L) out.insert(out.end(), T[j][kk].charges.begin(), T[j][kk].charges.end());
    else
      for (int c = 0; c < 4; ++c) {
        auto& ch = T[j + 1][4 * kk + c].in_chk;
        out.insert(out.end(), ch.begin(), ch.end());
      }
  };
  own_cols(k, boxA);
  for (int s = 0; s < 16; ++s)
    if (B.inner[s] != -1) il_rows(B.inner[s], IL);
  for (int s = 0; s < 24; ++s)
    if (B.outer[s] != -1) il_rows(B.outer[s], IL);
  std::vector<int> rb, cb;
  Mat Ac;
  aca(IL, boxA, rb, cb, Ac, B.Ar);
  B.out_chk.clear();
  B.out_chg.clear();
  if (!rb.empty()) {
    for (int r : rb) B.out_chk.push_back(IL[r]);
    for (int c : cb) B.out_chg.push_back(boxA[c]);
  }
}

// getNodes_incoming_box (FMM2DTree_find_error.hpp:1283-1336)
void Tree::nodes_incoming(int j, int k) {
  Box& B = T[j][k];
  std::vector<int> boxA, IL;
  auto own_rows = [&](int kk, std::vector<int>& out) {
    if (j == L) out.insert(out.end(), T[j][kk].charges.begin(), T[j][kk].charges.end());
    else
      for (int c = 0; c < 4; ++c) {
        auto& ch = T[j + 1][4 * kk + c].in_chk;
        out.insert(out.end(), ch.begin(), ch.end());
      }
  };
  auto il_cols = [&](int kk, std::vector<int>& out) {
    if (j == L) out.insert(out.end(), T[j][kk].charges.begin(), T[j][kk].charges.end());
    else
      for (int c = 0; c < 4; ++c) {
        auto& ch = T[j + 1][4 * kk + c].out_chg;
        out.insert(out.end(), ch.begin(), ch.end());
      }
  };
  own_rows(k, boxA);
  for (int s = 0; s < 16; ++s)
    if (B.inner[s] != -1) il_cols(B.inner[s], IL);
  for (int s = 0; s < 24; ++s)
    if (B.outer[s] != -1) il_cols(B.outer[s], IL);
  std::vector<int> rb, cb;
  Mat Ar;
  aca(boxA, IL, rb, cb, B.Ac, Ar);
  B.in_chk.clear();
  B.in_chg.clear();
  if (!rb.empty()) {
    for (int r : rb) B.in_chk.push_back(boxA[r]);
    for (int c : cb) B.in_chg.push_back(IL[c]);
  }
}

static Mat kmat(const Disc* D, const std::vector<int>& r, const std::vector<int>& c) {
  Mat A(r.size(), c.size());
  for (size_t j = 0; j < c.size(); ++j)
    for (size_t i = 0; i < r.size(); ++i) A(i, j) = D->entry(r[i], c[j]);
  return A;
}

// get_L2P_P2M_box (FMM2DTree_find_error.hpp:1105-1128)
void Tree::l2p_p2m(int j, int k) {
  Box& B = T[j][k];
  if (B.in_chk.size() < B.out_chk.size()) {
    B.out_chk.resize(B.in_chk.size());
    B.out_chg.resize(B.in_chk.size());
  } else {
    B.in_chk.resize(B.out_chk.size());
    B.in_chg.resize(B.out_chk.size());
  }
  int r = B.in_chk.size();
  Mat t1 = block(B.Ac, 0, 0, B.Ac.r, r);
  B.L2P = Mat(B.Ac.r, r);
  if (r > 0) {
    Mat D1 = kmat(D, B.in_chk, B.in_chg);
    LU lu;
    lu.compute(adj(D1));
    B.L2P = adj(lu.solve(adj(t1)));
  }
  Mat t2 = block(B.Ar, 0, 0, r, B.Ar.c);
  B.P2M = Mat(r, B.Ar.c);
  if (r > 0) {
    Mat D2 = kmat(D, B.out_chk, B.out_chg);
    LU lu;
    lu.compute(D2);
    B.P2M = lu.solve(t2);
  }
}

// getNodes (FMM2DTree_find_error.hpp:1137-1158); boxes of a level are
// independent, so they run in parallel here (the generator variant does the
// same, FMM2DTree_gen_rhs_x.hpp:1192,1201).
void Tree::getNodes() {
  for (int j = L; j >= 2; --j) {
    int nb = 1 << (2 * j);
#pragma omp parallel for schedule(dynamic)
    for (int k = 0; k < nb; ++k) nodes_outgoing(j, k);
#pragma omp parallel for schedule(dynamic)
    for (int k = 0; k < nb; ++k) nodes_incoming(j, k);
#pragma omp parallel for schedule(dynamic)
    for (int k = 0; k < nb; ++k) l2p_p2m(j, k);
  }
}

// assemble_M2L (FMM2DTree_find_error.hpp:1338-1361)
void Tree::assemble_M2L() {
  for (int j = 2; j <= L; ++j) {
    int nb = 1 << (2 * j);
#pragma omp parallel for schedule(dynamic)
    for (int k = 0; k < nb; ++k) {
      Box& B = T[j][k];
      for (int s = 0; s < 16; ++s)
        if (B.inner[s] != -1) B.M2L[B.inner[s]] = kmat(D, B.in_chk, T[j][B.inner[s]].out_chg);
      for (int s = 0; s < 24; ++s)
        if (B.outer[s] != -1) B.M2L[B.outer[s]] = kmat(D, B.in_chk, T[j][B.outer[s]].out_chg);
    }
  }
}

// initialise_phase + initialise_rhs (FMM2DTree_find_error.hpp:1703-1727)
void Tree::initialise_phase() {
  for (int j = 2; j <= L; ++j)
    for (auto& B : T[j]) {
      B.NumMultipoles = B.P2M.r;
      B.NumLocals = B.L2P.c;
      B.local_rhs = Mat(B.NumMultipoles, 1);
      B.multipole_rhs = Mat(B.NumLocals, 1);
    }
}

// initialise_P2P_Leaf_Level (FMM2DTree_find_error.hpp:1729-1741)
void Tree::initialise_P2P_leaf() {
  int nb = 1 << (2 * L);
#pragma omp parallel for schedule(dynamic)
  for (int k = 0; k < nb; ++k) {
    Box& B = T[L][k];
    for (int n = 0; n < 8; ++n)
      if (B.nb[n] != -1) B.P2P[B.nb[n]] = kmat(D, B.charges, T[L][B.nb[n]].charges);
    B.P2P[k] = kmat(D, B.charges, B.charges);
  }
}

// initialise_P2P_NonLeafLevel + assign_particle_rhs_NonLeafLevel
// (FMM2DTree_find_error.hpp:1743-1807)
void Tree::initialise_P2P_nonleaf(int j) {
  int nb = 1 << (2 * j);
  for (int k = 0; k < nb; ++k) {
    Box& B = T[j][k];
    std::vector<int> targets;
    for (int n = 0; n < 8; ++n)
      if (B.nb[n] != -1) targets.push_back(B.nb[n]);
    targets.push_back(k);
    for (int nn : targets) {
      int nr = 0, nc = 0;
      for (int a = 0; a < 4; ++a) nr += T[j + 1][4 * k + a].NumLocals;
      for (int b = 0; b < 4; ++b) nc += T[j + 1][4 * nn + b].NumMultipoles;
      Mat P(nr, nc);
      int ri = 0;
      for (int a = 0; a < 4; ++a) {
        Box& C = T[j + 1][4 * k + a];
        int ci = 0;
        for (int b = 0; b < 4; ++b) {
          int cc = 4 * nn + b;
          Mat& blk = C.M2L[cc];
          int br = C.NumLocals, bc = T[j + 1][cc].NumMultipoles;
          if (blk.r == br && blk.c == bc)
            for (int y = 0; y < bc; ++y)
              for (int x = 0; x < br; ++x) P(ri + x, ci + y) = blk(x, y);
          ci += bc;
        }
        ri += C.NumLocals;
      }
      B.P2P[nn] = P;
    }
    int len = 0;
    for (int c = 0; c < 4; ++c) len += T[j + 1][4 * k + c].NumLocals;
    B.particle_rhs = Mat(len, 1);
    int o = 0;
    for (int c = 0; c < 4; ++c) {
      Mat& mr = T[j + 1][4 * k + c].multipole_rhs;
      for (int i = 0; i < mr.r; ++i) B.particle_rhs(o + i, 0) = mr(i, 0);
      o += mr.r;
    }
  }
}

bool Tree::well_separated(int j, int k, int l) const {  // FMM2DTree_find_error.hpp:2077-2088
  if (k == l) return false;
  for (int i = 0; i < 8; ++i)
    if (T[j][k].nb[i] == l) return false;
  return true;
}

// ------------------------------------------------------------- eliminate
// eliminate_x (FMM2DTree_find_error.hpp:1814-1840, 1961-2075).  The solves
// with the self block are identical for identical operands, so each is done
// once per box.
void Tree::eliminate_x(int j, int k) {
  Box& K = T[j][k];
  Mat& A = K.P2P[k];
  bool have = A.size() != 0;
  CPQR qr;
  if (have) qr.compute(A);
  auto S = [&](const Mat& B) -> Mat { return have ? qr.solve(B) : Mat(A.c, B.c); };
  Mat WL = S(K.L2P);
  Mat Wb = S(K.particle_rhs);
  std::map<int, Mat> WP, WM;
  for (int n = 0; n < 8; ++n) {
    int nn = K.nb[n];
    if (nn == -1) continue;
    if (!T[j][nn].Eliminated) WP[nn] = S(K.P2P[nn]);
    else WM[nn] = S(K.M2P[nn]);
  }
  // fill-ins in k's P2M equation
  if (have) {
    K.L2M_f = neg(mul(K.P2M, WL));
    for (auto& e : WP) K.P2M_f[e.first] = neg(mul(K.P2M, e.second));
    for (auto& e : WM) K.M2M_f[e.first] = neg(mul(K.P2M, e.second));
    K.local_rhs = neg(mul(K.P2M, Wb));
  } else {
    K.L2M_f = Mat(K.P2M.r, K.L2P.c);
    for (auto& e : WP) K.P2M_f[e.first] = Mat(K.P2M.r, K.P2P[e.first].c);
    for (auto& e : WM) K.M2M_f[e.first] = Mat(K.P2M.r, K.M2P[e.first].c);
    K.local_rhs = Mat(K.P2M.r, 1);
  }
  for (int n = 0; n < 8; ++n) {
    int nn = K.nb[n];
    if (nn == -1) continue;
    Box& Nb = T[j][nn];
    bool ne = Nb.Eliminated;
    Mat C = ne ? Nb.P2L[k] : Nb.P2P[k];  // row block coupling nn to x_k
    Nb.L2P_f = have ? neg(mul(C, WL)) : Mat(C.r, K.L2P.c);
    for (int p = 0; p < 8; ++p) {
      int pp = K.nb[p];
      if (pp == -1) continue;
      bool pe = T[j][pp].Eliminated;
      if (!ne) {
        if (!pe) {
          Mat upd = have ? mul(C, WP[pp]) : Mat(C.r, K.P2P[pp].c);
          if (Nb.P2P[pp].size() == 0) Nb.P2P[pp] = neg(upd);
          else subeq(Nb.P2P[pp], upd);
          if (well_separated(j, nn, pp) && nn < pp && fro(Nb.P2P[pp]) > 1e-10) P2Plist.push_back({nn, pp});
        } else {
          Mat upd = have ? mul(C, WM[pp]) : Mat(C.r, K.M2P[pp].c);
          if (Nb.M2P[pp].size() == 0) Nb.M2P[pp] = neg(upd);
          else subeq(Nb.M2P[pp], upd);
        }
      } else {
        if (!pe) {
          Mat upd = have ? mul(C, WP[pp]) : Mat(C.r, K.P2P[pp].c);
          if (Nb.P2L[pp].size() == 0) Nb.P2L[pp] = neg(upd);
          else subeq(Nb.P2L[pp], upd);
          if (well_separated(j, nn, pp) && fro(Nb.P2L[pp]) > 1e-10) P2L_M2Plist.push_back({nn, pp});
        } else {
          if (have) subeq(Nb.M2L[pp], mul(C, WM[pp]));
        }
      }
    }
    if (have) {
      if (!ne) subeq(Nb.particle_rhs, mul(C, Wb));
      else subeq(Nb.multipole_rhs, mul(C, Wb));
    }
  }
}

// eliminate_z (FMM2DTree_find_error.hpp:2090-2113, 2245-2372)
void Tree::eliminate_z(int j, int k) {
  Box& K = T[j][k];
  bool have = K.L2M_f.size() > 0;
  CPQR qr;
  if (have) {
    qr.compute(K.L2M_f, piv_eps);
    if (qr.nonzero < std::min(qr.m, qr.n)) zero_piv.insert(zero_piv.end(), {j, k, qr.nonzero, std::min(qr.m, qr.n)});
  }
  int r = K.NumMultipoles;
  if (have) {
    K.M2L[k] = qr.solve(neg(eye(r)));
    for (int n = 0; n < 8; ++n) {
      int nn = K.nb[n];
      if (nn == -1) continue;
      if (!T[j][nn].Eliminated) K.P2L[nn] = qr.solve(K.P2M_f[nn]);
      else K.M2L[nn] = qr.solve(K.M2M_f[nn]);
    }
    K.multipole_rhs = qr.solve(K.local_rhs);
  } else {
    K.M2L[k] = Mat(0, r);
    for (int n = 0; n < 8; ++n) {
      int nn = K.nb[n];
      if (nn == -1) continue;
      if (!T[j][nn].Eliminated) K.P2L[nn] = Mat(0, K.P2M_f[nn].c);
      else K.M2L[nn] = Mat(0, K.M2M_f[nn].c);
    }
    K.multipole_rhs = Mat(0, 1);
  }
  Mat Xi = K.M2L[k];
  for (int n = 0; n < 8; ++n) {
    int nn = K.nb[n];
    if (nn == -1) continue;
    Box& Nb = T[j][nn];
    bool ne = Nb.Eliminated;
    // due to y
    Mat upd = have ? mul(Nb.L2P_f, Xi) : Mat(Nb.L2P_f.r, r);
    if (!ne) {
      if (Nb.M2P[k].size() == 0) Nb.M2P[k] = neg(upd);
      else subeq(Nb.M2P[k], upd);
    } else {
      if (Nb.M2L[k].size() == 0) Nb.M2L[k] = neg(upd);
      else subeq(Nb.M2L[k], upd);
    }
    // due to neighbours
    for (int p = 0; p < 8; ++p) {
      int pp = K.nb[p];
      if (pp == -1) continue;
      bool pe = T[j][pp].Eliminated;
      if (!ne) {
        if (!pe) {
          if (have) subeq(Nb.P2P[pp], mul(Nb.L2P_f, K.P2L[pp]));
          if (well_separated(j, nn, pp) && nn < pp && fro(Nb.P2P[pp]) > 1e-10) P2Plist.push_back({nn, pp});
        } else {
          if (have) subeq(Nb.M2P[pp], mul(Nb.L2P_f, K.M2L[pp]));
        }
      } else {
        if (!pe) {
          if (have) subeq(Nb.P2L[pp], mul(Nb.L2P_f, K.P2L[pp]));
          if (well_separated(j, nn, pp) && fro(Nb.P2L[pp]) > 1e-10) P2L_M2Plist.push_back({nn, pp});
        } else {
          if (have) subeq(Nb.M2L[pp], mul(Nb.L2P_f, K.M2L[pp]));
        }
      }
    }
    if (have) {
      if (!ne) subeq(Nb.particle_rhs, mul(Nb.L2P_f, K.multipole_rhs));
      else subeq(Nb.multipole_rhs, mul(Nb.L2P_f, K.multipole_rhs));
    }
  }
}

// ------------------------------------------------------------ compression
// compress_P2P (FMM2DTree_find_error.hpp:1532-1560)
void Tree::compress_P2P(int j, int k, int b, Mat& nL2P, Mat& nP2Mt, Mat& M2L) {
  Box& K = T[j][k];
  Box& B = T[j][b];
  Mat Acol = hcat(K.L2P, K.P2P[b]);
  CPQR q1;
  q1.compute(Acol);
  nL2P = q1.Q(q1.rank(rrqr_thr));
  Mat Rkp = mulAH(nL2P, K.P2P[b]);
  Mat Rk = mulAH(nL2P, K.L2P);
  Mat Arow = hcat(adj(B.P2M), adj(Rkp));
  CPQR q2;
  q2.compute(Arow);
  nP2Mt = q2.Q(q2.rank(rrqr_thr));
  Mat Rn = mulAH(nP2Mt, adj(B.P2M));
  Mat Rnp = mulAH(nP2Mt, adj(Rkp));
  M2L = mul(mul(Rk, K.M2L[b]), adj(Rn));
  Mat t = adj(Rnp);
  for (size_t i = 0; i < M2L.a.size(); ++i) M2L.a[i] += t.a[i];
}

// compress_M2P (FMM2DTree_find_error.hpp:1562-1575)
void Tree::compress_M2P(int j, int k, int b, Mat& nL2P, Mat& M2L) {
  Box& K = T[j][k];
  Mat Acol = hcat(K.L2P, K.M2P[b]);
  CPQR q1;
  q1.compute(Acol);
  nL2P = q1.Q(q1.rank(rrqr_thr));
  Mat Rkp = mulAH(nL2P, K.M2P[b]);
  Mat Rk = mulAH(nL2P, K.L2P);
  M2L = mul(Rk, K.M2L[b]);
  for (size_t i = 0; i < M2L.a.size(); ++i) M2L.a[i] += Rkp.a[i];
}

// compress_P2L (FMM2DTree_find_error.hpp:1577-1591)
void Tree::compress_P2L(int j, int k, int b, Mat& nP2Mt, Mat& M2L) {
  Box& K = T[j][k];
  Box& B = T[j][b];
  Mat Arow = hcat(adj(B.P2M), adj(K.P2L[b]));
  CPQR q2;
  q2.compute(Arow);
  nP2Mt = q2.Q(q2.rank(rrqr_thr));
  Mat Rn = mulAH(nP2Mt, adj(B.P2M));
  Mat Rnp = mulAH(nP2Mt, adj(K.P2L[b]));
  M2L = mul(K.M2L[b], adj(Rn));
  Mat t = adj(Rnp);
  for (size_t i = 0; i < M2L.a.size(); ++i) M2L.a[i] += t.a[i];
}

// update_other_M2Ls_at_k (FMM2DTree_find_error.hpp:1594-1629)
void Tree::update_M2L_at_k(int j, int k, int b, const Mat& nL2P) {
  Box& K = T[j][k];
  Mat R = mulAH(nL2P, K.L2P);
  for (int s = 0; s < 16; ++s) {
    int kil = K.inner[s];
    if (kil == -1 || kil == b) continue;
    K.M2L[kil] = mul(R, K.M2L[kil]);
  }
  for (int s = 0; s < 24; ++s) {
    int kil = K.outer[s];
    if (kil == -1 || kil == b) continue;
    K.M2L[kil] = mul(R, K.M2L[kil]);
  }
  for (int n = 0; n < 8; ++n) {
    int kn = K.nb[n];
    if (kn != -1 && T[j][kn].Eliminated && K.Eliminated) K.M2L[kn] = mul(R, K.M2L[kn]);
  }
  if (K.Eliminated) K.M2L[k] = mul(R, K.M2L[k]);
}

// update_other_M2Ls_at_b (FMM2DTree_find_error.hpp:1631-1669), including the
// outer-ring skip test on b (never true; fill-in pairs are always inner).
void Tree::update_M2L_at_b(int j, int k, int b, const Mat& nP2Mt) {
  Box& B = T[j][b];
  Mat R = mul(B.P2M, nP2Mt);
  for (int s = 0; s < 16; ++s) {
    int kil = B.inner[s];
    if (kil == -1 || kil == k) continue;
    T[j][kil].M2L[b] = mul(T[j][kil].M2L[b], R);
  }
  for (int s = 0; s < 24; ++s) {
    int kil = B.outer[s];
    if (kil == -1 || kil == b) continue;
    T[j][kil].M2L[b] = mul(T[j][kil].M2L[b], R);
  }
  for (int n = 0; n < 8; ++n) {
    int kn = B.nb[n];
    if (kn != -1 && T[j][kn].Eliminated && B.Eliminated) T[j][kn].M2L[b] = mul(T[j][kn].M2L[b], R);
  }
  if (B.Eliminated) B.M2L[b] = mul(B.M2L[b], R);
}

// update_parent_L2P / update_parent_P2M (FMM2DTree_find_error.hpp:1671-1699)
void Tree::update_parent_L2P(int j, int k, const Mat& nL2P) {
  int kP = k / 4, ck = k % 4;
  Mat& PL = T[j - 1][kP].L2P;
  Mat U[4];
  int ri = 0;
  for (int c = 0; c < 4; ++c) {
    int nl = T[j][kP * 4 + c].NumLocals;
    U[c] = block(PL, ri, 0, nl, PL.c);
    ri += nl;
  }
  U[ck] = mul(mulAH(nL2P, T[j][k].L2P), U[ck]);
  int tot = 0;
  for (int c = 0; c < 4; ++c) tot += U[c].r;
  Mat NP(tot, PL.c);
  ri = 0;
  for (int c = 0; c < 4; ++c) {
    for (int y = 0; y < PL.c; ++y)
      for (int x = 0; x < U[c].r; ++x) NP(ri + x, y) = U[c](x, y);
    ri += U[c].r;
  }
  PL = NP;
}

void Tree::update_parent_P2M(int j, int b, const Mat& nP2Mt) {
  int bP = b / 4, cb = b % 4;
  Mat& PM = T[j - 1][bP].P2M;
  Mat V[4];
  int ci = 0;
  for (int c = 0; c < 4; ++c) {
    int nm = T[j][bP * 4 + c].NumMultipoles;
    V[c] = block(PM, 0, ci, PM.r, nm);
    ci += nm;
  }
  V[cb] = mul(mul(V[cb], T[j][b].P2M), nP2Mt);
  int tot = 0;
  for (int c = 0; c < 4; ++c) tot += V[c].c;
  Mat NP(PM.r, tot);
  ci = 0;
  for (int c = 0; c < 4; ++c) {
    for (int y = 0; y < V[c].c; ++y)
      for (int x = 0; x < PM.r; ++x) NP(x, ci + y) = V[c](x, y);
    ci += V[c].c;
  }
  PM = NP;
}

static void equalize(const Mat& nL, const Mat& nPt, Mat& Lm, Mat& Pm) {
  if (nL.c < nPt.c) {
    Pm = block(nPt, 0, 0, nPt.r, nL.c);
    Lm = nL;
  } else {
    Lm = block(nL, 0, 0, nL.r, nPt.c);
    Pm = nPt;
  }
}

// compress_P2P_Two_Ways (FMM2DTree_find_error.hpp:1424-1490)
void Tree::compress_P2P_two_ways(int j, int k, int b) {
  Mat nLk, nPtb, M2Lkb, nLb, nPtk, M2Lbk;
  compress_P2P(j, k, b, nLk, nPtb, M2Lkb);
  compress_P2P(j, b, k, nLb, nPtk, M2Lbk);
  Mat Lk, Ptk, Lb, Ptb;
  equalize(nLk, nPtk, Lk, Ptk);
  equalize(nLb, nPtb, Lb, Ptb);
  Box& K = T[j][k];
  Box& B = T[j][b];
  K.M2L[b] = block(M2Lkb, 0, 0, Lk.c, Ptb.c);
  update_M2L_at_k(j, k, b, Lk);
  update_M2L_at_b(j, k, b, Ptb);
  if (j > 2) {
    update_parent_L2P(j, k, Lk);
    update_parent_P2M(j, b, Ptb);
  }
  K.L2P = Lk;
  B.P2M = adj(Ptb);
  K.NumLocals = K.L2P.c;
  B.NumMultipoles = B.P2M.r;
  B.M2L[k] = block(M2Lbk, 0, 0, Lb.c, Ptk.c);
  update_M2L_at_k(j, b, k, Lb);
  update_M2L_at_b(j, b, k, Ptk);
  if (j > 2) {
    update_parent_L2P(j, b, Lb);
    update_parent_P2M(j, k, Ptk);
  }
  B.L2P = Lb;
  K.P2M = adj(Ptk);
  B.NumLocals = B.L2P.c;
  K.NumMultipoles = K.P2M.r;
}

// compress_M2P_P2L (FMM2DTree_find_error.hpp:1492-1530); k un-eliminated,
// b eliminated.
void Tree::compress_M2P_P2L(int j, int k, int b) {
  Mat nLk, M2Lkb, nPtk, M2Lbk;
  compress_M2P(j, k, b, nLk, M2Lkb);
  compress_P2L(j, b, k, nPtk, M2Lbk);
  Mat Lk, Ptk;
  equalize(nLk, nPtk, Lk, Ptk);
  Box& K = T[j][k];
  Box& B = T[j][b];
  K.M2L[b] = block(M2Lkb, 0, 0, Lk.c, M2Lkb.c);
  update_M2L_at_k(j, k, b, Lk);
  if (j > 2) update_parent_L2P(j, k, Lk);
  K.L2P = Lk;
  K.NumLocals = K.L2P.c;
  B.M2L[k] = block(M2Lbk, 0, 0, M2Lbk.r, Ptk.c);
  update_M2L_at_b(j, b, k, Ptk);
  if (j > 2) update_parent_P2M(j, k, Ptk);
  K.P2M = adj(Ptk);
  K.NumMultipoles = K.P2M.r;
}

// Batched recompression used by the level-synchronous colour schedule (this
// build's GPU schedule; no reference counterpart).  After a colour step,
// every fill-in block between boxes at Chebyshev distance 2 whose Frobenius
// norm exceeds 1e-10 is folded into the bases: per un-eliminated box u,
// the row basis comes from CPQR([L2P_u | row fill-ins]) and the column
// basis from CPQR([P2M_u^H | column fill-ins^H]); the new rank is the smaller
// of the two thresholded ranks; M2L rows/columns of u and the parent's
// transfer blocks are rotated, and the fill-ins are projected onto M2L.
void Tree::colored_recompress(int j) {
  int nb = 1 << (2 * j);
  struct Fill {
    int u, p, kind;  // kind 0: P2P(u,p) 1: M2P(u,p) u unelim, p elim 2: P2L(u,p) u elim, p unelim
  };
  std::vector<Fill> fills;
  for (int u = 0; u < nb; ++u) {
    Box& U = T[j][u];
    for (int s = 0; s < 16; ++s) {
      int p = U.inner[s];
      if (p == -1) continue;
      bool ue = U.Eliminated, pe = T[j][p].Eliminated;
      Mat* blk = nullptr;
      int kind = -1;
      if (!ue && !pe) {
        auto it = U.P2P.find(p);
        if (it != U.P2P.end()) blk = &it->second, kind = 0;
      } else if (!ue && pe) {
        auto it = U.M2P.find(p);
        if (it != U.M2P.end()) blk = &it->second, kind = 1;
      } else if (ue && !pe) {
        auto it = U.P2L.find(p);
        if (it != U.P2L.end()) blk = &it->second, kind = 2;
      }
      if (!blk || blk->size() == 0) continue;
      if (fro(*blk) > 1e-10) fills.push_back({u, p, kind});
      else *blk = Mat();
    }
  }
  if (fills.empty()) return;
  std::vector<char> touched(nb, 0);
  std::vector<Mat> rowext(nb), colext(nb);
  for (auto& f : fills) {
    Box& U = T[j][f.u];
    if (f.kind == 0) {
      touched[f.u] = touched[f.p] = 1;
    } else if (f.kind == 1) {
      touched[f.u] = 1;
    } else {
      touched[f.p] = 1;
    }
    (void)U;
  }
  std::vector<Mat> QR_(nb), QC_(nb);
  std::vector<int> rnew(nb, 0);
  std::string dbg_line;
  int ntouch = 0;
  for (int u = 0; u < nb; ++u) ntouch += touched[u];
  for (int u = 0; u < nb; ++u) {
    if (!touched[u]) continue;
    Box& U = T[j][u];
    Mat Arow = U.L2P, Acol = adj(U.P2M);
    for (auto& f : fills) {
      if (f.u == u && f.kind == 0) Arow = hcat(Arow, T[j][u].P2P[f.p]);
      if (f.u == u && f.kind == 1) Arow = hcat(Arow, T[j][u].M2P[f.p]);
      if (f.p == u && f.kind == 0) Acol = hcat(Acol, adj(T[j][f.u].P2P[u]));
      if (f.p == u && f.kind == 2) Acol = hcat(Acol, adj(T[j][f.u].P2L[u]));
    }
    CPQR q1, q2;
    q1.compute(Arow);
    q2.compute(Acol);
    // equal row/column ranks, truncated to the smaller one as the reference
    // does (FMM2DTree_find_error.hpp:1445-1460, 1505-1512)
    int r = std::min(q1.rank(rrqr_thr), q2.rank(rrqr_thr));
    rnew[u] = r;
    rc_trace.insert(rc_trace.end(), {j, cur_color, u, q1.rank(rrqr_thr), q2.rank(rrqr_thr), r});
    if (getenv("AIFMM_DEBUG")) dbg_line += " " + std::to_string(u) + "(" + std::to_string(q1.rank(rrqr_thr)) + "," +
                                            std::to_string(q2.rank(rrqr_thr)) + "|r" + std::to_string(U.L2P.c) + ")";
    QR_[u] = q1.Q(r);
    QC_[u] = q2.Q(r);
  }
  if (getenv("AIFMM_DEBUG"))
    fprintf(stderr, "[orc] level %d: %zu fills, %d touched:%s\n", j, fills.size(), ntouch, dbg_line.c_str());
  // rotate M2L rows (row basis change) then columns (column basis change)
  std::vector<Mat> Rrow(nb), Rcol(nb);
  for (int u = 0; u < nb; ++u)
    if (touched[u]) {
      Rrow[u] = mulAH(QR_[u], T[j][u].L2P);   // rnew x r_old
      Rcol[u] = mul(T[j][u].P2M, QC_[u]);     // r_old x rnew
    }
  for (int u = 0; u < nb; ++u) {
    if (!touched[u]) continue;
    Box& U = T[j][u];
    for (int s = 0; s < 16; ++s)
      if (U.inner[s] != -1) U.M2L[U.inner[s]] = mul(Rrow[u], U.M2L[U.inner[s]]);
    for (int s = 0; s < 24; ++s)
      if (U.outer[s] != -1) U.M2L[U.outer[s]] = mul(Rrow[u], U.M2L[U.outer[s]]);
  }
  for (int u = 0; u < nb; ++u) {
    if (!touched[u]) continue;
    Box& U = T[j][u];
    for (int s = 0; s < 16; ++s)
      if (U.inner[s] != -1) T[j][U.inner[s]].M2L[u] = mul(T[j][U.inner[s]].M2L[u], Rcol[u]);
    for (int s = 0; s < 24; ++s)
      if (U.outer[s] != -1) T[j][U.outer[s]].M2L[u] = mul(T[j][U.outer[s]].M2L[u], Rcol[u]);
  }
  for (auto& f : fills) {
    Box& U = T[j][f.u];
    Mat add;
    if (f.kind == 0) add = mul(mulAH(QR_[f.u], U.P2P[f.p]), QC_[f.p]);
    if (f.kind == 1) add = mulAH(QR_[f.u], U.M2P[f.p]);
    if (f.kind == 2) add = mul(U.P2L[f.p], QC_[f.p]);
    Mat& M = U.M2L[f.p];
    for (size_t i = 0; i < M.a.size(); ++i) M.a[i] += add.a[i];
    if (f.kind == 0) U.P2P[f.p] = Mat();
    if (f.kind == 1) U.M2P[f.p] = Mat();
    if (f.kind == 2) U.P2L[f.p] = Mat();
  }
  // parent transfer blocks (child row/column slices), then the new bases
  if (j > 2) {
    for (int P = 0; P < nb / 4; ++P) {
      bool any = false;
      for (int c = 0; c < 4; ++c) any |= touched[4 * P + c] != 0;
      if (!any) continue;
      Mat& PL = T[j - 1][P].L2P;
      Mat& PM = T[j - 1][P].P2M;
      int totr = 0, totc = 0;
      for (int c = 0; c < 4; ++c) {
        int ch = 4 * P + c;
        totr += touched[ch] ? rnew[ch] : T[j][ch].NumLocals;
        totc += touched[ch] ? rnew[ch] : T[j][ch].NumMultipoles;
      }
      Mat NL(totr, PL.c), NM(PM.r, totc);
      int ri = 0, ro = 0, ci = 0, co = 0;
      for (int c = 0; c < 4; ++c) {
        int ch = 4 * P + c;
        int nl = T[j][ch].NumLocals, nm = T[j][ch].NumMultipoles;
        Mat Ub = block(PL, ri, 0, nl, PL.c);
        Mat Vb = block(PM, 0, ci, PM.r, nm);
        if (touched[ch]) {
          Ub = mul(Rrow[ch], Ub);
          Vb = mul(Vb, Rcol[ch]);
        }
        for (int y = 0; y < PL.c; ++y)
          for (int x = 0; x < Ub.r; ++x) NL(ro + x, y) = Ub(x, y);
        for (int y = 0; y < Vb.c; ++y)
          for (int x = 0; x < PM.r; ++x) NM(x, co + y) = Vb(x, y);
        ri += nl, ro += Ub.r, ci += nm, co += Vb.c;
      }
      PL = NL;
      PM = NM;
    }
  }
  for (int u = 0; u < nb; ++u) {
    if (!touched[u]) continue;
    Box& U = T[j][u];
    U.L2P = QR_[u];
    U.P2M = adj(QC_[u]);
    U.NumLocals = U.NumMultipoles = rnew[u];
    U.local_rhs = Mat(rnew[u], 1);
    U.multipole_rhs = Mat(rnew[u], 1);
  }
}

// eliminate_phase_efficient / eliminate_phase (FMM2DTree_find_error.hpp:2374-2451)
void Tree::eliminate(Schedule s) {
  for (int j = L; j >= 2; --j) {
    if (j != L) initialise_P2P_nonleaf(j);
    int nb = 1 << (2 * j);
    order[j].clear();
    if (s == SCHED_COLORED) {
      for (int c = 0; c < 9; ++c) {
        for (int k = 0; k < nb; ++k) {
          int ix, iy;
          box_xy(j, k, ix, iy);
          if ((ix % 3) + 3 * (iy % 3) != c) continue;
          T[j][k].Eliminated = true;
          eliminate_x(j, k);
          eliminate_z(j, k);
          order[j].push_back(k);
        }
        P2Plist.clear();
        P2L_M2Plist.clear();
        cur_color = c;
        colored_recompress(j);
      }
      continue;
    }
    for (int k = 0; k < nb; ++k) {
      if (s == SCHED_EFFICIENT) {
        for (size_t i = 0; i < P2Plist.size(); i++) {
          int nn = P2Plist[i].first, pp = P2Plist[i].second;
          if (k == nn || k == pp) {
            P2Plist.erase(P2Plist.begin() + i);
            i--;
            if (T[j][nn].P2P[pp].size() != 0) {
              compress_P2P_two_ways(j, nn, pp);
              T[j][nn].P2P[pp] = Mat();
              T[j][pp].P2P[nn] = Mat();
            }
          }
        }
        for (size_t i = 0; i < P2L_M2Plist.size(); i++) {
          int nn = P2L_M2Plist[i].first, pp = P2L_M2Plist[i].second;
          if (k == nn || k == pp) {
            P2L_M2Plist.erase(P2L_M2Plist.begin() + i);
            i--;
            if (T[j][nn].P2L[pp].size() != 0) {
              compress_M2P_P2L(j, pp, nn);
              T[j][nn].P2L[pp] = Mat();
              T[j][pp].M2P[nn] = Mat();
            }
          }
        }
        P2Plist.clear();
        P2L_M2Plist.clear();
        T[j][k].Eliminated = true;
        eliminate_x(j, k);
        eliminate_z(j, k);
      } else {
        T[j][k].Eliminated = true;
        eliminate_x(j, k);
        eliminate_z(j, k);
        // duplicates (pushed by both the x and the z step) are skipped once
        // their block has been cleared.
        for (auto& pr : P2Plist) {
          int nn = pr.first, pp = pr.second;
          if (T[j][nn].P2P[pp].size() == 0) continue;
          compress_P2P_two_ways(j, nn, pp);
          T[j][nn].P2P[pp] = Mat();
          T[j][pp].P2P[nn] = Mat();
        }
        P2Plist.clear();
        for (auto& pr : P2L_M2Plist) {
          int nn = pr.first, pp = pr.second;
          if (T[j][nn].P2L[pp].size() == 0) continue;
          compress_M2P_P2L(j, pp, nn);
          T[j][nn].P2L[pp] = Mat();
          T[j][pp].M2P[nn] = Mat();
        }
        P2L_M2Plist.clear();
      }
      order[j].push_back(k);
    }
  }
  base_solve();
}

// solve_particles_at_base_level (FMM2DTree_find_error.hpp:2453-2485)
void Tree::base_solve() {
  int j = 2, nb = 16;
  int NM = 0, NL = 0;
  for (int k = 0; k < nb; ++k) NM += T[j][k].NumMultipoles, NL += T[j][k].NumLocals;
  if (NL == 0 || NM == 0) return;
  Mat A(NL, NM), b(NL, 1);
  int ri = 0;
  for (int r = 0; r < nb; ++r) {
    int si = 0;
    for (int s = 0; s < nb; ++s) {
      Mat& blk = T[j][r].M2L[s];
      if (blk.r == T[j][r].NumLocals && blk.c == T[j][s].NumMultipoles)
        for (int y = 0; y < blk.c; ++y)
          for (int x = 0; x < blk.r; ++x) A(ri + x, si + y) = blk(x, y);
      si += T[j][s].NumMultipoles;
    }
    Mat& mr = T[j][r].multipole_rhs;
    for (int x = 0; x < mr.r; ++x) b(ri + x, 0) = mr(x, 0);
    ri += T[j][r].NumLocals;
  }
  CPQR qr;
  qr.compute(A, base_piv_eps);
  double mx = 0, mn = HUGE_VAL;
  for (int i = 0; i < std::min(qr.m, qr.n); ++i) mx = std::max(mx, std::abs(qr.qr(i, i))), mn = std::min(mn, std::abs(qr.qr(i, i)));
  base_rdiag_ratio = mx > 0 ? mn / mx : 0;
  Mat y = qr.solve(b);
  int idx = 0;
  for (int k = 0; k < nb; ++k) {
    int nm = T[j][k].NumMultipoles;
    T[j][k].multipoles = block(y, idx, 0, nm, 1);
    idx += nm;
  }
}

// back_substitution_phase (FMM2DTree_find_error.hpp:2487-2532), boxes in
// reverse elimination order.
void Tree::back_substitution() {
  for (int j = 2; j <= L; ++j) {
    for (int oi = (int)order[j].size() - 1; oi >= 0; --oi) {
      int k = order[j][oi];
      Box& K = T[j][k];
      if (K.P2P[k].size() != 0) {
        CPQR qr;
        qr.compute(K.P2P[k]);
        Mat rhs1 = K.particle_rhs;
        for (int n = 0; n < 8; ++n) {
          int nn = K.nb[n];
          if (nn == -1) continue;
          if (!T[j][nn].Eliminated) subeq(rhs1, mul(K.P2P[nn], T[j][nn].particles));
          else subeq(rhs1, mul(K.M2P[nn], T[j][nn].multipoles));
        }
        Mat rhs2 = K.multipoles;
        if (rhs2.r == 0 && K.P2M.r == 0) rhs2 = Mat(0, 1);
        subeq(rhs2, mul(K.P2M, qr.solve(rhs1)));
        Mat Atemp = neg(mul(K.P2M, qr.solve(K.L2P)));
        if (Atemp.size() != 0) {
          CPQR q2;
          q2.compute(Atemp, piv_eps);
          K.locals = q2.solve(rhs2);
        } else {
          K.locals = Mat(0, 1);
        }
        Mat t = rhs1;
        subeq(t, mul(K.L2P, K.locals));
        K.particles = qr.solve(t);
      } else {
        K.particles = Mat(0, 1);
      }
      K.Eliminated = false;
      if (j != L) {
        int idx = 0;
        for (int c = 0; c < 4; ++c) {
          int nm = T[j + 1][4 * k + c].NumMultipoles;
          T[j + 1][4 * k + c].multipoles = block(K.particles, idx, 0, nm, 1);
          idx += nm;
        }
      }
    }
  }
}

void Tree::solution(cd* x) const {
  int r = D->rank;
  for (int k = 0; k < (1 << (2 * L)); ++k) {
    const Mat& p = T[L][k].particles;
    for (int i = 0; i < r; ++i) x[(int64_t)k * r + i] = (p.r == r) ? p(i, 0) : cd(0, 0);
  }
}

}  // namespace orc
