// Level-synchronous colour-batched elimination (9 colour classes per level:
// boxes with equal (ix mod 3, iy mod 3) have disjoint 3x3 neighbourhoods, so
// their eliminations commute and write disjoint blocks), followed after each
// colour by batched recompression of the distance-2 fill-ins.
//
// Per box k (reference: eliminate_x / eliminate_z,
// FMM2DTree_find_error.hpp:1814-2372), with A = P2P[k][k]:
//   W   = A^{-1} [L2P | E1 row blocks of k]             (eliminate x_k)
//   G   = -P2M W ;  L2M_f = G[:, :r]                       (fill-ins of E2_k)
//   X   = L2M_f^{-1} [-I | G[:, r:]]                       (eliminate z_k)
//         -> M2L[k][k], P2L[k][nn] / M2L[k][nn]            (new E3_k)
//   V   = [0 | W[:, r:]] - W[:, :r] X
//   every neighbour equation nn:  row(nn) -= C_nn V  with C_nn = P2P[nn][k]
//   (nn not yet eliminated) or P2L[nn][k] (nn eliminated).
#include "factor.h"

#include <algorithm>
#include <array>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <type_traits>
#include <unordered_map>

namespace aifmm {

static double now_s() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}
static const double2 ONE = {1.0, 0.0}, MONE = {-1.0, 0.0}, ZERO = {0.0, 0.0};

// Zero-pivot epsilon of the L2M_f factorisations (complex128, Factor::l2mf_eps,
// 1e-12 by default).  Eigen's ColPivHouseholderQR counts a pivot as nonzero
// down to 2.2e-16 of the largest column norm; an L2M_f whose trailing pivot
// sits below 1e-12 of its leading one (a low-contrast box whose two skeleton
// directions coincide to rounding) has no correct digit left in that pivot,
// and its inverse would carry that rounding into every fill-in by up to 1e15.
// Such pivots are treated as zero: the solve is Eigen's basic solution, as for
// an exactly singular block.

template <class T>
Factor<T>::Factor(Problem& p) : pb(p), L(p.L), R(p.R) {
  ar.set_chunk(Arena::chunk_for(p.N, sizeof(T) * (2 << 10), 32 << 20, 1 << 30));
  sc.set_chunk(Arena::chunk_for(p.N, 2 << 10, 16 << 20, 256 << 20));
  sv.set_chunk(Arena::chunk_for(p.N, 2 << 10, 16 << 20, 256 << 20));
  pv.set_chunk(Arena::chunk_for(p.N, 1 << 10, 16 << 20, 256 << 20));
  spare_.pv.set_chunk(Arena::chunk_for(p.N, 1 << 10, 16 << 20, 256 << 20));
}

template <class T>
size_t Factor<T>::device_bytes() const {
  return ar.reserved() + sc.reserved() + sv.reserved() + pv.reserved() + spare_.pv.reserved();
}

template <class T>
void Factor<T>::init_level(int j) {
  int nb = 1 << (2 * j);
  B[j].assign(nb, FBox());
  colors[j].assign(9, {});
  CvtBatch<T> cv;
  auto conv = [&](const Blk<double2>& s) {
    Blk<T> d = new_blk<T>(ar, s.r, s.c);
    cv.add(s.p, d.p, (int64_t)s.size());
    return d;
  };
  for (int k = 0; k < nb; ++k) {
    FBox& F = B[j][k];
    const SkelBox& S = pb.sk[j][k];
    F.m = S.m;
    F.r = S.r;
    int ix, iy;
    box_xy(j, k, ix, iy);
    F.color = (ix % 3) + 3 * (iy % 3);
    colors[j][F.color].push_back(k);
    F.L2P = conv(S.L2P);
    F.P2M = conv(S.P2M);
    // (sources in key order: each insertion goes at the end)
    for (auto& e : S.M2L) F.M2L.emplace_hint(F.M2L.end(), e.first, conv(e.second));
    if (j == L)
      for (auto& e : pb.P2P[k]) F.P2P.emplace_hint(F.P2P.end(), e.first, conv(e.second));
  }
  cv.run(sc, pb.stream);
  sc.reset();  // scratch reuse is ordered on the stream
}

// Parent P2P blocks from the children's M2L blocks
// (initialise_P2P_NonLeafLevel, FMM2DTree_find_error.hpp:1743-1807).
template <class T>
void Factor<T>::nonleaf_p2p(int j) {
  int nb = 1 << (2 * j);
  CopyBatch<T> cp, zero;
  for (int k = 0; k < nb; ++k) {
    FBox& K = B[j][k];
    int mk = 0;
    for (int a = 0; a < 4; ++a) mk += B[j + 1][4 * k + a].r;
    if (mk != K.m || K.L2P.r != mk)
      throw std::runtime_error("level " + std::to_string(j) + " box " + std::to_string(k) + ": size mismatch");
    std::vector<int> tg;
    for (int n = 0; n < 8; ++n)
      if (pb.lists[j][k].nb[n] != -1) tg.push_back(pb.lists[j][k].nb[n]);
    tg.push_back(k);
    for (int nn : tg) {
      Blk<T> P = new_blk<T>(ar, K.m, B[j][nn].m);
      zero.add(P, P, 4);
      int ri = 0;
      for (int a = 0; a < 4; ++a) {
        FBox& C = B[j + 1][4 * k + a];
        int ci = 0;
        for (int b = 0; b < 4; ++b) {
          int cc = 4 * nn + b;
          int bc = B[j + 1][cc].r;
          auto it = C.M2L.find(cc);
          if (it != C.M2L.end() && it->second.r == C.r && it->second.c == bc && C.r && bc)
            cp.add(it->second, P.sub(ri, ci, C.r, bc));
          ci += bc;
        }
        ri += C.r;
      }
      K.P2P[nn] = P;
    }
  }
  zero.run(sc, pb.stream);
  cp.run(sc, pb.stream);
  sc.reset();  // scratch reuse is ordered on the stream
}

template <class T>
void Factor<T>::eliminate_color(int j, int c) {
  auto& Bx = B[j];
  const auto& LS = pb.lists[j];
  const std::vector<int>& S = colors[j][c];
  std::vector<std::array<int, 8>> woff(S.size());
  std::vector<int> ncols(S.size());
  CopyBatch<T> cp;
  DenseOps<T> dn;
  GemmBatch<T> g;
  double fl = 0;
  double h0 = now_s();
  // 1. W = [L2P | row blocks], LU of A, W <- A^{-1} W
  for (size_t t = 0; t < S.size(); ++t) {
    int k = S[t];
    FBox& K = Bx[k];
    int m = K.m, r = K.r;
    int off = r;
    for (int n = 0; n < 8; ++n) {
      int nn = LS[k].nb[n];
      woff[t][n] = off;
      if (nn < 0) continue;
      off += Bx[nn].elim ? Bx[nn].r : Bx[nn].m;
    }
    ncols[t] = off;
    K.W = new_blk<T>(ar, m, off);
    if (m == 0) continue;
    cp.add(K.L2P, K.W.sub(0, 0, m, r));
    for (int n = 0; n < 8; ++n) {
      int nn = LS[k].nb[n];
      if (nn < 0) continue;
      int w = Bx[nn].elim ? Bx[nn].r : Bx[nn].m;
      if (!w) continue;
      auto& mp = Bx[nn].elim ? K.M2P : K.P2P;
      auto it = mp.find(nn);
      Blk<T> dst = K.W.sub(0, woff[t][n], m, w);
      if (it != mp.end() && it->second.r == m && it->second.c == w) cp.add(it->second, dst);
      else cp.add(dst, dst, 4);
    }
    K.A = K.P2P[k];
    dn.factor(ar, sc, K.fa, K.A, 0, gate_thr);
    fl += 16.0 * m * m * m / 3.0 + 16.0 * m * m * off;
  }
  cp.run(sc, pb.stream);
  dn.run_factor(sc, pb.stream);
  for (size_t t = 0; t < S.size(); ++t) {
    FBox& K = Bx[S[t]];
    if (K.m) dn.solve(sc, K.fa, K.W);
  }
  dn.run_solve(sc, pb.stream);
  hs_[0] += now_s() - h0, h0 = now_s();
  // 2. G = -P2M W
  for (size_t t = 0; t < S.size(); ++t) {
    FBox& K = Bx[S[t]];
    K.G = new_blk<T>(ar, K.r, ncols[t]);
    if (K.m && K.r) {
      g.add(K.P2M, 0, K.W, 0, K.G, MONE, ZERO);
      fl += 8.0 * K.r * K.m * ncols[t];
    }
  }
  g.run(sc, pb.stream);
  // 3. LU of L2M_f ; 4. X = L2M_f^{-1} [-I | G[:, r:]]
  std::vector<Blk<T>> X(S.size());
  for (size_t t = 0; t < S.size(); ++t) {
    FBox& K = Bx[S[t]];
    int r = K.r;
    X[t] = new_blk<T>(ar, r, ncols[t]);
    if (!r) continue;
    dn.factor(ar, sc, K.fg, K.G.sub(0, 0, r, r), l2mf_eps, gate_thr);
    cp.add(X[t], X[t].sub(0, 0, r, r), 3, MONE);
    if (ncols[t] > r) cp.add(K.G.sub(0, r, r, ncols[t] - r), X[t].sub(0, r, r, ncols[t] - r));
    fl += 16.0 * r * r * r / 3.0 + 16.0 * r * r * ncols[t];
  }
  dn.run_factor(sc, pb.stream);
  cp.run(sc, pb.stream);
  for (size_t t = 0; t < S.size(); ++t) {
    FBox& K = Bx[S[t]];
    if (K.r) dn.solve(sc, K.fg, X[t]);
  }
  dn.run_solve(sc, pb.stream);
  // 5. V = [0 | W[:, r:]] - W[:, :r] X
  std::vector<Blk<T>> V(S.size());
  for (size_t t = 0; t < S.size(); ++t) {
    int k = S[t];
    FBox& K = Bx[k];
    int m = K.m, r = K.r, nc = ncols[t];
    V[t] = new_blk<T>(ar, m, nc);
    // new E3_k blocks (views into X)
    K.M2L[k] = X[t].sub(0, 0, r, r);
    for (int n = 0; n < 8; ++n) {
      int nn = LS[k].nb[n];
      if (nn < 0) continue;
      if (Bx[nn].elim) K.M2L[nn] = X[t].sub(0, woff[t][n], r, Bx[nn].r);
      else K.P2L[nn] = X[t].sub(0, woff[t][n], r, Bx[nn].m);
    }
    if (!m) continue;
    if (r) cp.add(V[t], V[t].sub(0, 0, m, r), 4);
    if (nc > r) cp.add(K.W.sub(0, r, m, nc - r), V[t].sub(0, r, m, nc - r));
  }
  cp.run(sc, pb.stream);
  for (size_t t = 0; t < S.size(); ++t) {
    FBox& K = Bx[S[t]];
    if (K.m && K.r) {
      g.add(K.W.sub(0, 0, K.m, K.r), 0, X[t], 0, V[t], MONE, ONE);
      fl += 8.0 * K.m * K.r * ncols[t];
    }
  }
  g.run(sc, pb.stream);
  hs_[1] += now_s() - h0, h0 = now_s();
  // 6. neighbour equations: target -= C_nn V[:, block]
  CopyBatch<T> zero;
  auto ensure = [&](std::map<int, Blk<T>>& mp, int key, int rows, int cols) -> Blk<T> {
    auto it = mp.find(key);
    if (it != mp.end() && it->second.r == rows && it->second.c == cols && (it->second.p || !rows || !cols))
      return it->second;
    if (it != mp.end() && it->second.size())
      throw std::runtime_error("block size mismatch in neighbour update");
    Blk<T> b = new_blk<T>(ar, rows, cols);
    zero.add(b, b, 4);
    mp[key] = b;
    return b;
  };
  for (size_t t = 0; t < S.size(); ++t) {
    int k = S[t];
    FBox& K = Bx[k];
    int m = K.m, r = K.r;
    for (int n = 0; n < 8; ++n) {
      int nn = LS[k].nb[n];
      if (nn < 0) continue;
      FBox& Nb = Bx[nn];
      bool ne = Nb.elim;
      int rows = ne ? Nb.r : Nb.m;
      auto& cm = ne ? Nb.P2L : Nb.P2P;
      Blk<T> C;
      auto it = cm.find(k);
      if (it != cm.end()) C = it->second;
      if (!rows || !m || C.size() == 0) {
        // still create the structural blocks so later steps find them
        ensure(ne ? Nb.M2L : Nb.M2P, k, rows, r);
        for (int p = 0; p < 8; ++p) {
          int pp = LS[k].nb[p];
          if (pp < 0) continue;
          bool pe = Bx[pp].elim;
          auto& tm = ne ? (pe ? Nb.M2L : Nb.P2L) : (pe ? Nb.M2P : Nb.P2P);
          ensure(tm, pp, rows, pe ? Bx[pp].r : Bx[pp].m);
        }
        continue;
      }
      Blk<T> ty = ensure(ne ? Nb.M2L : Nb.M2P, k, rows, r);
      if (r) g.add(C, 0, V[t].sub(0, 0, m, r), 0, ty, MONE, ONE), fl += 8.0 * rows * m * r;
      for (int p = 0; p < 8; ++p) {
        int pp = LS[k].nb[p];
        if (pp < 0) continue;
        bool pe = Bx[pp].elim;
        int cols = pe ? Bx[pp].r : Bx[pp].m;
        auto& tm = ne ? (pe ? Nb.M2L : Nb.P2L) : (pe ? Nb.M2P : Nb.P2P);
        Blk<T> tg = ensure(tm, pp, rows, cols);
        if (cols) g.add(C, 0, V[t].sub(0, woff[t][p], m, cols), 0, tg, MONE, ONE), fl += 8.0 * rows * m * cols;
      }
    }
  }
  zero.run(sc, pb.stream);
  g.run(sc, pb.stream);
  hs_[2] += now_s() - h0;
  for (int k : S) Bx[k].elim = true;
  sc.reset();  // scratch reuse is ordered on the stream
  st.flops += fl;
}

// Batched recompression of distance-2 fill-ins (the colour-schedule form of
// compress_P2P_Two_Ways / compress_M2P_P2L, FMM2DTree_find_error.hpp:1424-1699):
// per un-eliminated box u, row basis = CPQR([L2P_u | row fill-ins]), column
// basis = CPQR([P2M_u^H | column fill-ins^H]); rank = min of the two
// thresholded ranks (the reference's equalisation).
template <class T>
void Factor<T>::recompress(int j, int c) {
  using Rl = typename CplxOf<T>::real;
  auto& Bx = B[j];
  const auto& LS = pb.lists[j];
  int nb = 1 << (2 * j);
  struct Fill {
    int u, p, kind;
  };
  double h0 = now_s();
  std::vector<Fill> cand;
  std::vector<const Blk<T>*> cb;
  auto blk_of = [&](const Fill& f) -> Blk<T>& {
    FBox& U = Bx[f.u];
    return f.kind == 0 ? U.P2P[f.p] : f.kind == 1 ? U.M2P[f.p] : U.P2L[f.p];
  };
  for (int u = 0; u < nb; ++u) {
    FBox& U = Bx[u];
    for (int s = 0; s < 16; ++s) {
      int p = LS[u].inner[s];
      if (p < 0) continue;
      bool ue = U.elim, pe = Bx[p].elim;
      std::map<int, Blk<T>>* mp = nullptr;
      int kind = -1;
      if (!ue && !pe) mp = &U.P2P, kind = 0;
      else if (!ue && pe) mp = &U.M2P, kind = 1;
      else if (ue && !pe) mp = &U.P2L, kind = 2;
      if (!mp) continue;
      auto it = mp->find(p);
      if (it == mp->end()) continue;
      if (it->second.size() == 0) {
        mp->erase(it);
        continue;
      }
      cand.push_back(Fill{u, p, kind});
      cb.push_back(&it->second);
    }
  }
  // (inside a merged run every member takes part in each launch round, so
  // an early return comes after the rounds of this function's first launch)
  FroBatch<T> fro;
  double* dn = cand.empty() ? nullptr : sc.get<double>(cand.size());
  for (size_t i = 0; i < cand.size(); ++i) fro.add(cb[i]->p, cb[i]->r, cb[i]->c, cb[i]->ld, dn + i);
  fro.run(sc, pb.stream);
  hs_[3] += now_s() - h0;
  if (cand.empty()) return;
  std::vector<double> norms(cand.size());
  {
    const double w0 = now_s();
    d2h(norms.data(), dn, 8 * cand.size(), pb.stream);
    t_wait_ += now_s() - w0;
  }
  h0 = now_s();
  std::vector<Fill> fills;
  for (size_t i = 0; i < cand.size(); ++i) {
    if (norms[i] > 1e-10) fills.push_back(cand[i]);
    else {
      FBox& U = Bx[cand[i].u];
      (cand[i].kind == 0 ? U.P2P : cand[i].kind == 1 ? U.M2P : U.P2L).erase(cand[i].p);
    }
  }
  if (fills.empty()) {
    sc.reset();
    return;
  }
  std::vector<char> touched(nb, 0);
  for (auto& f : fills) {
    if (f.kind == 0) touched[f.u] = touched[f.p] = 1;
    else if (f.kind == 1) touched[f.u] = 1;
    else touched[f.p] = 1;
  }
  std::vector<int> tl;
  for (int u = 0; u < nb; ++u)
    if (touched[u]) tl.push_back(u);
  int nt = (int)tl.size();
  st.recompressions += nt;
  std::vector<int> ti(nb, -1);
  for (int i = 0; i < nt; ++i) ti[tl[i]] = i;
  // each touched box's row fills (it is the target) and column fills (it is
  // the source), in fill order: the column order of its two stacks
  std::vector<std::vector<int>> rowf(nt), colf(nt);
  for (size_t q = 0; q < fills.size(); ++q) {
    const Fill& f = fills[q];
    if (f.kind == 0 || f.kind == 1) rowf[ti[f.u]].push_back((int)q);
    if (f.kind == 0 || f.kind == 2) colf[ti[f.p]].push_back((int)q);
  }
  // stacked matrices
  std::vector<Blk<T>> Arow(nt), Acol(nt);
  CopyBatch<T> cp;
  for (int i = 0; i < nt; ++i) {
    int u = tl[i];
    FBox& U = Bx[u];
    int m = U.m;
    int nr = U.r, ncl = U.r;
    for (int q : rowf[i]) nr += fills[q].kind == 0 ? Bx[fills[q].p].m : Bx[fills[q].p].r;
    for (int q : colf[i]) ncl += fills[q].kind == 0 ? Bx[fills[q].u].m : Bx[fills[q].u].r;
    Arow[i] = new_blk<T>(ar, m, nr);
    Acol[i] = new_blk<T>(ar, m, ncl);
    if (!m) continue;
    int oR = 0, oC = 0;
    if (U.r) cp.add(U.L2P, Arow[i].sub(0, 0, m, U.r)), cp.add(U.P2M, Acol[i].sub(0, 0, m, U.r), 1);
    oR = oC = U.r;
    for (int q : rowf[i]) {
      Blk<T>& F = blk_of(fills[q]);
      if (F.c) cp.add(F, Arow[i].sub(0, oR, m, F.c));
      oR += F.c;
    }
    for (int q : colf[i]) {
      Blk<T>& F = blk_of(fills[q]);
      if (F.r) cp.add(F, Acol[i].sub(0, oC, m, F.r), 1);
      oC += F.r;
    }
  }
  cp.run(sc, pb.stream);
  // CPQR of both stacks
  QRBatch<T> qb;
  std::vector<QRDesc<T>>& qd = qb.d;
  int* dr = sc.get<int>(2 * nt);
  std::vector<T*> taus(2 * nt);
  for (int i = 0; i < 2 * nt; ++i) {
    Blk<T>& A = (i % 2 == 0) ? Arow[i / 2] : Acol[i / 2];
    QRDesc<T> d{};
    d.A = A.p;
    d.m = A.r;
    d.n = A.c;
    d.lda = A.ld;
    int mn = std::min(A.r, A.c);
    taus[i] = ar.get<T>(mn > 0 ? mn : 1);
    d.tau = taus[i];
    d.perm = sc.get<int>(A.c > 0 ? A.c : 1);
    d.work = sc.get<Rl>(A.c > 0 ? 2 * A.c : 2);
    d.thr = pow(10.0, -1.0 * pb.P.tol_pow);
    d.rank_out = dr + i;
    d.maxpiv_out = nullptr;
    d.early = 1;
    if (A.r && A.c) qd.push_back(d);
    else HIP_OK(hipMemsetAsync(dr + i, 0, 4, pb.stream));
    double mm = A.r, nn = A.c;
    st.flops += mm >= nn ? 8.0 * (mm * nn * nn - nn * nn * nn / 3) : 8.0 * (nn * mm * mm - mm * mm * mm / 3);
  }
  qb.run(sc, pb.stream);
  hs_[4] += now_s() - h0;
  std::vector<int> rk(2 * nt);
  const double w1 = now_s();
  d2h(rk.data(), dr, 8 * nt, pb.stream);
  t_wait_ += now_s() - w1;
  std::vector<int> rnew(nb, -1);
  for (int i = 0; i < nt; ++i) {
    const int u = tl[i];
    const int r = Bx[u].m ? std::min(rk[2 * i], rk[2 * i + 1]) : 0;
    st.rc_trace.insert(st.rc_trace.end(), {j, c, u, rk[2 * i], rk[2 * i + 1], r});
  }
  static const bool dbg = getenv("AIFMM_DEBUG") != nullptr;
  if (dbg) {
    fprintf(stderr, "[gpu] level %d: %zu candidates, %zu fills, %d touched:", j, cand.size(), fills.size(), nt);
    for (int i = 0; i < nt; ++i)
      fprintf(stderr, " %d(%d,%d|r%d|%dx%d,%d)", tl[i], rk[2 * i], rk[2 * i + 1], Bx[tl[i]].r, Arow[i].r, Arow[i].c,
              Acol[i].c);
    fprintf(stderr, "\n");
  }
  h0 = now_s();
  std::vector<Blk<T>> QR(nt), QC(nt), Rrow(nt), Rcol(nt);
  QFormBatch<T> qfb;
  std::vector<QFormDesc<T>>& qf = qfb.d;
  GemmBatch<T> g;
  for (int i = 0; i < nt; ++i) {
    int u = tl[i];
    FBox& U = Bx[u];
    int r = std::min(rk[2 * i], rk[2 * i + 1]);
    if (!U.m) r = 0;
    rnew[u] = r;
    QR[i] = new_blk<T>(ar, U.m, r);
    QC[i] = new_blk<T>(ar, U.m, r);
    if (r) {
      qf.push_back(QFormDesc<T>{Arow[i].p, taus[2 * i], QR[i].p, U.m, Arow[i].c, r, Arow[i].ld, QR[i].ld});
      qf.push_back(QFormDesc<T>{Acol[i].p, taus[2 * i + 1], QC[i].p, U.m, Acol[i].c, r, Acol[i].ld, QC[i].ld});
      st.flops += 2 * 8.0 * U.m * r * std::min(U.m, std::max(Arow[i].c, Acol[i].c));
    }
    Rrow[i] = new_blk<T>(ar, r, U.r);
    Rcol[i] = new_blk<T>(ar, U.r, r);
  }
  qfb.run(sc, pb.stream);
  for (int i = 0; i < nt; ++i) {
    FBox& U = Bx[tl[i]];
    if (rnew[tl[i]] && U.r) {
      g.add(QR[i], 1, U.L2P, 0, Rrow[i], ONE, ZERO);
      g.add(U.P2M, 0, QC[i], 0, Rcol[i], ONE, ZERO);
    }
  }
  g.run(sc, pb.stream);
  auto il_of = [&](int u, std::vector<int>& v) {
    v.clear();
    for (int s = 0; s < 16; ++s)
      if (LS[u].inner[s] != -1) v.push_back(LS[u].inner[s]);
    for (int s = 0; s < 24; ++s)
      if (LS[u].outer[s] != -1) v.push_back(LS[u].outer[s]);
  };
  hs_[5] += now_s() - h0, h0 = now_s();
  std::vector<int> il;
  // row rotation of M2L[u][*]
  for (int i = 0; i < nt; ++i) {
    int u = tl[i];
    FBox& U = Bx[u];
    il_of(u, il);
    for (int b : il) {
      Blk<T>& M = U.M2L[b];
      Blk<T> nw = new_blk<T>(ar, rnew[u], M.c);
      if (nw.size()) g.add(Rrow[i], 0, M, 0, nw, ONE, ZERO), st.flops += 8.0 * nw.r * nw.c * U.r;
      M = nw;
    }
  }
  g.run(sc, pb.stream);
  // column rotation of M2L[*][u]
  for (int i = 0; i < nt; ++i) {
    int u = tl[i];
    FBox& U = Bx[u];
    il_of(u, il);
    for (int b : il) {
      Blk<T>& M = Bx[b].M2L[u];
      Blk<T> nw = new_blk<T>(ar, M.r, rnew[u]);
      if (nw.size()) g.add(M, 0, Rcol[i], 0, nw, ONE, ZERO), st.flops += 8.0 * nw.r * nw.c * U.r;
      M = nw;
    }
  }
  g.run(sc, pb.stream);
  hs_[6] += now_s() - h0, h0 = now_s();
  // fill-ins projected onto M2L
  std::vector<Blk<T>> tmp(fills.size());
  for (size_t q = 0; q < fills.size(); ++q) {
    auto& f = fills[q];
    if (f.kind != 0) continue;
    int iu = ti[f.u];
    Blk<T>& F = blk_of(f);
    tmp[q] = new_blk<T>(ar, rnew[f.u], F.c);
    if (tmp[q].size()) g.add(QR[iu], 1, F, 0, tmp[q], ONE, ZERO);
  }
  g.run(sc, pb.stream);
  for (size_t q = 0; q < fills.size(); ++q) {
    auto& f = fills[q];
    Blk<T>& M = Bx[f.u].M2L[f.p];
    Blk<T>& F = blk_of(f);
    if (!M.size()) continue;
    if (f.kind == 0) g.add(tmp[q], 0, QC[ti[f.p]], 0, M, ONE, ONE);
    if (f.kind == 1) g.add(QR[ti[f.u]], 1, F, 0, M, ONE, ONE);
    if (f.kind == 2) g.add(F, 0, QC[ti[f.p]], 0, M, ONE, ONE);
  }
  g.run(sc, pb.stream);
  // parent transfer blocks
  if (j > 2) {
    for (int P = 0; P < nb / 4; ++P) {
      bool any = false;
      for (int c = 0; c < 4; ++c) any |= touched[4 * P + c] != 0;
      if (!any) continue;
      FBox& Pb = B[j - 1][P];
      int totr = 0;
      for (int c = 0; c < 4; ++c) totr += touched[4 * P + c] ? rnew[4 * P + c] : Bx[4 * P + c].r;
      Blk<T> NL = new_blk<T>(ar, totr, Pb.L2P.c), NM = new_blk<T>(ar, Pb.P2M.r, totr);
      int ri = 0, ro = 0;
      for (int c = 0; c < 4; ++c) {
        int ch = 4 * P + c;
        int ol = Bx[ch].r, nl = touched[ch] ? rnew[ch] : ol;
        if (touched[ch]) {
          int i = ti[ch];
          if (NL.c && nl) g.add(Rrow[i], 0, Pb.L2P.sub(ri, 0, ol, Pb.L2P.c), 0, NL.sub(ro, 0, nl, NL.c), ONE, ZERO);
          if (NM.r && nl) g.add(Pb.P2M.sub(0, ri, Pb.P2M.r, ol), 0, Rcol[i], 0, NM.sub(0, ro, NM.r, nl), ONE, ZERO);
        } else {
          if (NL.c && nl) cp.add(Pb.L2P.sub(ri, 0, ol, Pb.L2P.c), NL.sub(ro, 0, nl, NL.c));
          if (NM.r && nl) cp.add(Pb.P2M.sub(0, ri, Pb.P2M.r, ol), NM.sub(0, ro, NM.r, nl));
        }
        ri += ol;
        ro += nl;
      }
      Pb.L2P = NL;
      Pb.P2M = NM;
      Pb.m = totr;
    }
    g.run(sc, pb.stream);
    cp.run(sc, pb.stream);
  }
  // new bases
  for (int i = 0; i < nt; ++i) {
    int u = tl[i];
    FBox& U = Bx[u];
    U.L2P = QR[i];
    U.P2M = new_blk<T>(ar, rnew[u], U.m);
    if (U.P2M.size()) cp.add(QC[i], U.P2M, 1);
    U.r = rnew[u];
  }
  cp.run(sc, pb.stream);
  for (auto& f : fills) {
    FBox& U = Bx[f.u];
    (f.kind == 0 ? U.P2P : f.kind == 1 ? U.M2P : U.P2L).erase(f.p);
  }
  hs_[7] += now_s() - h0;
  sc.reset();  // scratch reuse is ordered on the stream
}

// solve_particles_at_base_level (FMM2DTree_find_error.hpp:2453-2485): one
// dense LU of the level-2 system.
template <class T>
void Factor<T>::base_factor() {
  auto& Bx = B[2];
  boff.assign(17, 0);
  for (int k = 0; k < 16; ++k) boff[k + 1] = boff[k] + Bx[k].r;
  int n = boff[16];
  st.base_n = n;
  base = new_blk<T>(ar, n, n);

  if (!n) return;
  CopyBatch<T> cp;
  cp.add(base, base, 4);
  cp.run(sc, pb.stream);
  for (int r = 0; r < 16; ++r)
    for (int s = 0; s < 16; ++s) {
      auto it = Bx[r].M2L.find(s);
      if (it == Bx[r].M2L.end()) continue;
      if (it->second.r == Bx[r].r && it->second.c == Bx[s].r && it->second.size())
        cp.add(it->second, base.sub(boff[r], boff[s], Bx[r].r, Bx[s].r));
    }
  cp.run(sc, pb.stream);
  DenseOps<T> dn;
  dn.factor(ar, sc, fb, base, base_eps);
  dn.run_factor(sc, pb.stream);
  st.flops += 16.0 * n * (double)n * n / 3.0;
  {
    // smallest over largest |diagonal| of the factor (reported in the stats)
    Blk<double2> f = DenseOps<T>::fp64_factor(fb);
    Blk<double2> dd = new_blk<double2>(sc, n, 1);
    CopyBatch<double2> dg;
    dg.add(Blk<double2>{f.p, 1, n, f.ld + 1}, dd, 2);
    dg.run(sc, pb.stream);
    std::vector<double2> hd(n);
    d2h(hd.data(), dd.p, n * sizeof(double2), pb.stream);
    HIP_OK(hipStreamSynchronize(pb.stream));
    double mx = 0, mn = HUGE_VAL;
    for (auto& v : hd) mx = std::max(mx, std::hypot(v.x, v.y)), mn = std::min(mn, std::hypot(v.x, v.y));
    st.base_rdiag_ratio = mx > 0 ? mn / mx : 0;
  }
  sc.reset();  // scratch reuse is ordered on the stream
}

template <class T>
void Factor<T>::factor() {
  if (!pb.assembled) pb.assemble();
  double t0 = now_s();
  B.assign(L + 1, {});
  colors.assign(L + 1, {});
  // AIFMM_FACTOR_TIMERS=1: host wall time per phase on stderr (a phase's time
  // includes waits for the device where it synchronises)
  static const bool timers = getenv("AIFMM_FACTOR_TIMERS") != nullptr;
  double tph[8] = {0};
  auto tick = [&](int k, double& t) {
    const double n = now_s();
    tph[k] += n - t;
    t = n;
  };
  double tt = now_s();
  for (int j = 2; j <= L; ++j) init_level(j);
  tick(0, tt);
  for (int j = L; j >= 2; --j) {
    if (j != L) nonleaf_p2p(j);
    tick(1, tt);
    for (int c = 0; c < 9; ++c) {
      if (colors[j][c].empty()) continue;
      eliminate_color(j, c);
      tick(2, tt);
      recompress(j, c);
      tick(3, tt);
    }
  }
  base_factor();
  tick(4, tt);
  if constexpr (std::is_same<T, double2>::value) {
    // nonzero-pivot counts of every L2M_f factorisation (one gather, one readback)
    std::vector<const int*> src;
    std::vector<std::array<int, 3>> who;
    for (int j = 2; j <= L; ++j)
      for (int k = 0; k < (int)B[j].size(); ++k)
        if (B[j][k].m && B[j][k].r) src.push_back(B[j][k].fg.q.nz), who.push_back({j, k, B[j][k].r});
    if (!src.empty()) {
      auto* ds = upload(sc, src, pb.stream);
      int* dn = sc.get<int>(src.size());
      launch_gather_i32(ds, dn, (int)src.size(), pb.stream);
      std::vector<int> nz(src.size());
      d2h(nz.data(), dn, nz.size() * 4, pb.stream);
      HIP_OK(hipStreamSynchronize(pb.stream));
      for (size_t i = 0; i < nz.size(); ++i)
        if (nz[i] < who[i][2]) st.zero_piv.insert(st.zero_piv.end(), {who[i][0], who[i][1], nz[i], who[i][2]});
    }
  } else if (gate_thr >= 0) {
    // which pivot blocks kept their complex64 LU (one gather, one readback)
    std::vector<const int*> src;
    for (int j = 2; j <= L; ++j)
      for (auto& F : B[j]) {
        if (F.fa.flag) src.push_back(F.fa.flag);
        if (F.fg.flag) src.push_back(F.fg.flag);
      }
    st.piv_c64 = st.piv_fp64 = 0;
    if (!src.empty()) {
      auto* ds = upload(sc, src, pb.stream);
      int* dn = sc.get<int>(src.size());
      launch_gather_i32(ds, dn, (int)src.size(), pb.stream);
      std::vector<int> fl(src.size());
      d2h(fl.data(), dn, fl.size() * 4, pb.stream);
      for (int v : fl) (v ? st.piv_fp64 : st.piv_c64) += 1;
    }
    for (int j = 2; j <= L; ++j)
      for (auto& F : B[j]) {
        if (F.m && !F.fa.flag) ++st.piv_fp64;
        if (F.m && F.r && !F.fg.flag) ++st.piv_fp64;
      }
  }
  tick(4, tt);
  double h0 = now_s();
  sweep_operators();
  hs_[8] += now_s() - h0, h0 = now_s();
  compose_sweep();
  hs_[9] += now_s() - h0;
  tick(5, tt);
  dense_top_alloc();
  h0 = now_s();
  build_plan(1);
  hs_[10] += now_s() - h0;
  tick(6, tt);
  dense_top_fill();
  HIP_OK(hipStreamSynchronize(pb.stream));
  tick(7, tt);
  sc.release();  // factorisation scratch; later users allocate afresh
  st.t_factor = now_s() - t0;
  if (timers)
    fprintf(stderr,
            "[factor] init %.1f ms, nonleaf P2P %.1f, eliminate %.1f, recompress %.1f, base %.1f, sweep operators "
            "%.1f, plan %.1f, dense top %.1f; recompress waits %.1f ms\n",
            1e3 * tph[0], 1e3 * tph[1], 1e3 * tph[2], 1e3 * tph[3], 1e3 * tph[4], 1e3 * tph[5], 1e3 * tph[6],
            1e3 * tph[7], 1e3 * t_wait_);
  if (timers)
    fprintf(stderr,
            "[factor] host: eliminate W/LU %.1f ms, G/X/V %.1f, neighbour updates %.1f; recompress scan %.1f, "
            "stacks %.1f, Q/R %.1f, M2L rotations %.1f, fills/parents/bases %.1f; sweep operators %.1f, compose "
            "%.1f, plan %.1f\n",
            1e3 * hs_[0], 1e3 * hs_[1], 1e3 * hs_[2], 1e3 * hs_[3], 1e3 * hs_[4], 1e3 * hs_[5], 1e3 * hs_[6],
            1e3 * hs_[7], 1e3 * hs_[8], 1e3 * hs_[9], 1e3 * hs_[10]);
  st.bytes = ar.used();
  for (int j = 2; j <= L; ++j)
    for (auto& F : B[j]) st.max_rank = std::max(st.max_rank, F.r);
}

// Sweep operators of every eliminated box (the products the forward sweep
// and back-substitution apply per box, formed once):
//   Z = A^{-1},  H = -P2M Z,  H <- L2M_f^{-1} H,  Z <- Z - W_r H,
//   U = W_r L2M_f^{-1},  and the inverse of the level-2 system.
template <class T>
void Factor<T>::sweep_operators() {
  CopyBatch<T> cp;
  DenseOps<T> svb;
  GemmBatch<T> g;
  std::vector<Blk<T>> ginv;
  // condition estimate of every L2M_f from the diagonal of its factor
  // (max |d_i| / min |d_i|): one strided copy of the diagonals, one readback
  {
    CopyBatch<double2> dg;
    size_t nd = 0;
    for (int j = 2; j <= L; ++j)
      for (auto& F : B[j])
        if (F.m && F.r) nd += F.r;
    Blk<double2> dd = new_blk<double2>(sc, (int)nd, 1);
    size_t o = 0;
    for (int j = 2; j <= L; ++j)
      for (auto& F : B[j]) {
        if (!F.m || !F.r) continue;
        Blk<double2> f = DenseOps<T>::fp64_factor(F.fg);
        Blk<double2> diag{f.p, 1, F.r, f.ld + 1};  // element (0, i) is f(i, i)
        dg.add(diag, dd.sub((int)o, 0, F.r, 1), 2);
        o += F.r;
      }
    dg.run(sc, pb.stream);
    std::vector<double2> hd(nd);
    if (nd) d2h(hd.data(), dd.p, nd * sizeof(double2), pb.stream);
    HIP_OK(hipStreamSynchronize(pb.stream));
    o = 0;
    st.ill_boxes = 0;
    for (int j = 2; j <= L; ++j)
      for (auto& F : B[j]) {
        F.ill = false;
        if (!F.m || !F.r) continue;
        double mx = 0, mn = HUGE_VAL;
        for (int i = 0; i < F.r; ++i) {
          const double a = std::hypot(hd[o + i].x, hd[o + i].y);
          mx = std::max(mx, a);
          mn = std::min(mn, a);
        }
        o += F.r;
        F.ill = !(mx <= kIllCond * mn);
        st.ill_boxes += F.ill;
      }
  }
  int rmax = 0;
  for (int j = 2; j <= L; ++j)
    for (auto& F : B[j]) {
      F.Z = new_blk<T>(ar, F.m, F.m);
      F.H = new_blk<T>(ar, F.r, F.m);
      F.U = new_blk<T>(ar, F.m, F.r);
      if (F.m) cp.add(F.Z, F.Z, 3);
      if (F.ill) {
        F.Ai = new_blk<T>(ar, F.m, F.m);
        F.PA = new_blk<T>(ar, F.r, F.m);
        rmax = std::max(rmax, F.r);
      }
    }
  eye = new_blk<T>(ar, rmax, rmax);
  if (rmax) cp.add(eye, eye, 3);
  cp.run(sc, pb.stream);
  for (int j = 2; j <= L; ++j)
    for (auto& F : B[j])
      if (F.m) svb.solve(sc, F.fa, F.Z), st.flops += 8.0 * F.m * F.m * F.m;
  svb.run_solve(sc, pb.stream);
  for (int j = 2; j <= L; ++j)
    for (auto& F : B[j])
      if (F.m && F.r) g.add(F.P2M, 0, F.Z, 0, F.H, MONE, ZERO), st.flops += 8.0 * F.r * F.m * F.m;
  g.run(sc, pb.stream);
  for (int j = 2; j <= L; ++j)
    for (auto& F : B[j])
      if (F.ill) cp.add(F.Z, F.Ai), cp.add(F.H, F.PA);
  cp.run(sc, pb.stream);
  for (int j = 2; j <= L; ++j)
    for (auto& F : B[j])
      if (F.m && F.r) svb.solve(sc, F.fg, F.H), st.flops += 8.0 * F.r * F.r * F.m;
  svb.run_solve(sc, pb.stream);
  for (int j = 2; j <= L; ++j)
    for (auto& F : B[j])
      if (F.m && F.r) g.add(F.W.sub(0, 0, F.m, F.r), 0, F.H, 0, F.Z, MONE, ONE), st.flops += 8.0 * F.m * F.r * F.m;
  g.run(sc, pb.stream);
  // U = W_r L2M_f^{-1} via the r x r inverse
  for (int j = 2; j <= L; ++j)
    for (auto& F : B[j]) {
      Blk<T> gi = new_blk<T>(sc, F.r, F.r);
      ginv.push_back(gi);
      if (F.m && F.r) cp.add(gi, gi, 3);
    }
  cp.run(sc, pb.stream);
  size_t q = 0;
  for (int j = 2; j <= L; ++j)
    for (auto& F : B[j]) {
      if (F.m && F.r) svb.solve(sc, F.fg, ginv[q]), st.flops += 8.0 * F.r * F.r * F.r;
      ++q;
    }
  svb.run_solve(sc, pb.stream);
  q = 0;
  for (int j = 2; j <= L; ++j)
    for (auto& F : B[j]) {
      if (F.m && F.r) g.add(F.W.sub(0, 0, F.m, F.r), 0, ginv[q], 0, F.U, ONE, ZERO), st.flops += 8.0 * F.m * F.r * F.r;
      ++q;
    }
  g.run(sc, pb.stream);
  int n = boff.empty() ? 0 : boff[16];
  binv = new_blk<T>(ar, n, n);
  if (n) {
    cp.add(binv, binv, 3);
    cp.run(sc, pb.stream);
    svb.solve(sc, fb, binv);
    svb.run_solve(sc, pb.stream);
    st.flops += 8.0 * n * (double)n * n;
  }
  sc.reset();  // scratch reuse is ordered on the stream
}

// Couplings of the sweeps composed with the sweep operator of the box they
// read from, formed once per factorisation:
//   P2P[k][nn] <- P2P[k][nn] Z_nn             nn regular
//   PU[k][nn]   = P2P[k][nn] U_nn             nn regular, colour nn > colour k
//   P2L[nn][k] <- P2L[nn][k] Z_k              k regular, colour k > colour nn
// (ill boxes keep their couplings: their outputs are formed in the
// reference's order and read as they are).  Every P2P[k][nn] is read by one
// sweep only (the forward when nn comes earlier, the back-substitution when
// later), so the product replaces it.  The products are formed in FP64 from
// the stored factors and rounded once to the factor's precision.
namespace {
template <class T>
struct Composer {
  // one product C = A B into dst (dst may be A)
  struct Job {
    Blk<T> a, b, dst;
  };
  std::vector<Job> jobs;
  size_t bytes = 0;  // scratch the pending jobs take
  void add(const Blk<T>& a, const Blk<T>& b, const Blk<T>& dst) {
    jobs.push_back(Job{a, b, dst});
    bytes += sizeof(double2) * ((size_t)a.size() + b.size() + dst.size());
  }
  void run(Arena& sc, hipStream_t s) {
    if (jobs.empty()) return;
    GemmBatch<double2> g;
    if constexpr (std::is_same<T, double2>::value) {
      CopyBatch<double2> cp;
      for (auto& j : jobs) {
        Blk<double2> t = new_blk<double2>(sc, j.dst.r, j.dst.c);
        g.add(j.a, 0, j.b, 0, t, ONE, ZERO);
        cp.add(t, j.dst);
      }
      g.run(sc, s);
      cp.run(sc, s);
    } else {
      WidenBatch wd;
      CvtBatch<float2> nd;
      std::unordered_map<const float2*, Blk<double2>> wide;  // each source widened once
      wide.reserve(2 * jobs.size());
      auto widen = [&](const Blk<float2>& x) {
        if (x.ld != x.r) throw std::runtime_error("sweep composition needs contiguous blocks");
        auto it = wide.find(x.p);
        if (it != wide.end()) return it->second;
        Blk<double2> w = new_blk<double2>(sc, x.r, x.c);
        wd.add(x.p, w.p, (int64_t)x.size());
        wide[x.p] = w;
        return w;
      };
      for (auto& j : jobs) {
        Blk<double2> a = widen(j.a), b = widen(j.b);
        Blk<double2> t = new_blk<double2>(sc, j.dst.r, j.dst.c);
        g.add(a, 0, b, 0, t, ONE, ZERO);
        if (j.dst.ld != j.dst.r) throw std::runtime_error("sweep composition needs contiguous blocks");
        nd.add(t.p, j.dst.p, (int64_t)j.dst.size());
      }
      wd.run(sc, s);
      g.run(sc, s);  // every product reads the widened couplings before any is replaced
      nd.run(sc, s);
    }
    jobs.clear();
    bytes = 0;
    sc.reset();  // scratch reuse is ordered on the stream
  }
};
}  // namespace

template <class T>
void Factor<T>::compose_sweep() {
  const size_t chunk = size_t(256) << 20;
  Composer<T> C;
  for (int j = 2; j <= L; ++j) {
    const auto& LS = pb.lists[j];
    for (int k = 0; k < (1 << (2 * j)); ++k) {
      FBox& F = B[j][k];
      for (int n = 0; n < 8; ++n) {
        const int nn = LS[k].nb[n];
        if (nn < 0) continue;
        FBox& Fn = B[j][nn];
        if (Fn.ill || !Fn.m) continue;
        auto it = F.P2P.find(nn);
        if (F.m && it != F.P2P.end() && it->second.size()) {
          Blk<T>& P = it->second;
          if (Fn.color > F.color && Fn.r) {
            Blk<T> pu = new_blk<T>(ar, P.r, Fn.r);
            C.add(P, Fn.U, pu);
            st.flops += 8.0 * P.r * (double)P.c * Fn.r;
            F.PU[nn] = pu;
          }
          C.add(P, Fn.Z, P);
          st.flops += 8.0 * P.r * (double)P.c * P.c;
        }
        // P2L held by box k towards a later-coloured regular neighbour nn
        auto il = F.P2L.find(nn);
        if (F.r && Fn.color > F.color && il != F.P2L.end() && il->second.size()) {
          C.add(il->second, Fn.Z, il->second);
          st.flops += 8.0 * il->second.r * (double)il->second.c * il->second.c;
        }
      }
      if (C.bytes > chunk) C.run(sc, pb.stream);
    }
  }
  C.run(sc, pb.stream);
}

// The forward sweep (the RHS updates the reference fuses into eliminate_*),
// the base solve and the back-substitution (back_substitution_phase,
// FMM2DTree_find_error.hpp:2487-2532) as one box program.  With y_k the
// phase-1 vector of box k and the composed couplings above:
//  forward, level j = L..2:
//     colour c = 0..8, box k of colour c:
//        y_k = b_k - sum_{nn earlier colour} P2P[k][nn] y_nn   (= P2P Z_nn y_nn)
//     (the eliminated b_nn = Z_nn y_nn is never formed), then every box nn:
//        mr_nn = H_nn y_nn - sum_{k later colour} P2L[nn][k] y_k
//     (mr_nn = rows of the parent's b)
//  base: xb = (level-2 system)^{-1} yb       (yb = the level-2 mr stacked)
//  backward, level j = 2..L:
//     colour c = 8..0, box k of colour c:
//        y_k = bs_k - sum_{nn later colour} (P2P[k][nn] y_nn - PU[k][nn] yp_nn)
//                   - sum_{nn earlier colour} M2P[k][nn] yp_nn
//     (yp_nn = rows of the parent's x, known before the level starts), then
//     every box:  x_k = Z_k y_k - U_k yp_k
// A colour's launch reads only phase-1 vectors of earlier launches, so each
// colour is one launch even when its boxes are cut into row tiles, and each
// level's outputs are one more.  An ill box (reference order) forms its
// outputs right after its colour, and its neighbours read those:
//     forward   mr_k = H y_k ;  b_k = A^{-1} y_k - W_r mr_k
//     backward  v_k = yp_k - P2M A^{-1} y_k ;  x_k = A^{-1} y_k - U v_k
// A pulled update equals the reference's pushed one: each neighbour update
// is applied before the value it changes is next read.
template <class T>
static void need_shape(const Blk<T>& b, int rows, int cols, const char* what) {
  if (b.r != rows || b.c != cols) throw std::runtime_error(std::string("sweep block shape mismatch: ") + what);
}

// Vectors of the sweep at levels 2..jhi for cap right-hand sides, from arena
// a.  The leaves (jhi == L) are slices of Bt / Xt.  With b2 / x2 given, the
// boxes of the dense level jd have their b and x as slices of them (n2 x cap,
// boxes stacked in order), so the sweep at levels jd..2 can run as one dense
// product.
template <class T>
void Factor<T>::alloc_vectors(Arena& a, int cap, int jhi, const Blk<double2>* b2, const Blk<double2>* x2) {
  if (jhi == L) {
    Bt = new_blk<double2>(a, (int)pb.N, cap);
    Xt = new_blk<double2>(a, (int)pb.N, cap);
  }
  const int nbase = boff.empty() ? 0 : boff[16];
  yb = new_blk<double2>(a, nbase, cap);
  xb = new_blk<double2>(a, nbase, cap);
  bk.assign(L + 1, {});
  bs.assign(L + 1, {});
  mr.assign(L + 1, {});
  xk.assign(L + 1, {});
  yk.assign(L + 1, {});
  yt.assign(L + 1, {});
  vk.assign(L + 1, {});
  for (int j = jhi; j >= 2; --j) {
    int nb = 1 << (2 * j);
    bk[j].resize(nb);
    bs[j].resize(nb);
    xk[j].resize(nb);
    for (int k = 0; k < nb; ++k) {
      FBox& F = B[j][k];
      if (j == L) {
        if (F.m != R) throw std::runtime_error("leaf block size mismatch");
        bk[j][k] = Bt.sub(k * R, 0, R, cap);
        xk[j][k] = Xt.sub(k * R, 0, R, cap);
      } else if (j == jd && b2) {
        bk[j][k] = b2->sub(off2[k], 0, F.m, cap);
        xk[j][k] = x2->sub(off2[k], 0, F.m, cap);
      } else {
        bk[j][k] = new_blk<double2>(a, F.m, cap);
        xk[j][k] = new_blk<double2>(a, F.m, cap);
      }
      bs[j][k] = new_blk<double2>(a, F.m, cap);
    }
  }
  for (int j = jhi; j >= 2; --j) {
    int nb = 1 << (2 * j);
    mr[j].resize(nb);
    yk[j].resize(nb);
    for (int k = 0; k < nb; ++k) {
      int r = B[j][k].r;
      if (j == 2) {
        mr[j][k] = yb.sub(boff[k], 0, r, cap);
        yk[j][k] = xb.sub(boff[k], 0, r, cap);
      } else {
        int o = 0;
        for (int c = 0; c < (k & 3); ++c) o += B[j][(k & ~3) + c].r;
        mr[j][k] = bk[j - 1][k >> 2].sub(o, 0, r, cap);
        yk[j][k] = xk[j - 1][k >> 2].sub(o, 0, r, cap);
      }
    }
  }
  // phase-1 vectors of tall boxes in the back-substitution (no save target)
  for (int j = 2; j <= jhi; ++j) {
    yt[j].resize(1 << (2 * j));
    vk[j].resize(1 << (2 * j));
    for (int k = 0; k < (1 << (2 * j)); ++k) {
      if (B[j][k].m) yt[j][k] = new_blk<double2>(a, B[j][k].m, cap);
      if (B[j][k].ill) vk[j][k] = new_blk<double2>(a, B[j][k].r, cap);
    }
  }
}

static int env_int(const char* name, int dflt) {
  const char* e = getenv(name);
  return e && *e ? atoi(e) : dflt;
}
static int tile_colour() { static const int v = env_int("AIFMM_TILE_COLOUR", 0); return v; }
static int tile_out() { static const int v = env_int("AIFMM_TILE_OUT", 64); return v; }

// forward sweep of levels jhi..jlo
template <class T>
void Factor<T>::emit_forward(Program<T>& P, int jlo, int jhi) {
  for (int j = jhi; j >= jlo; --j) {
    const auto& LS = pb.lists[j];
    for (int c = 0; c < 9; ++c) {
      P.stage();
      for (int k : colors[j][c]) {
        FBox& F = B[j][k];
        if (!F.m) continue;
        BoxOp<T> op;
        op.m = F.m;
        op.src = bk[j][k].p;
        op.ldsrc = bk[j][k].ld;
        op.save = bs[j][k].p;
        op.ldsave = bs[j][k].ld;
        op.tile = tile_colour();
        for (int n = 0; n < 8; ++n) {
          int nn = LS[k].nb[n];
          if (nn < 0 || B[j][nn].color >= c || !B[j][nn].m) continue;
          auto it = F.P2P.find(nn);
          if (it == F.P2P.end() || !it->second.size()) continue;
          need_shape(it->second, F.m, B[j][nn].m, "P2P (forward)");
          const Blk<double2>& v = B[j][nn].ill ? bk[j][nn] : bs[j][nn];
          op.term(it->second.p, it->second.ld, it->second.c, v.p, v.ld, MONE);
        }
        P.add(op);
      }
      // ill boxes of this colour: mr_k = H y ; b_k = A^{-1} y - W_r mr_k
      bool staged = false;
      for (int k : colors[j][c]) {
        FBox& F = B[j][k];
        if (!F.ill || !F.m) continue;
        if (!staged) P.stage(), staged = true;
        BoxOp<T> op;
        op.m = F.r;
        op.save = mr[j][k].p;
        op.ldsave = mr[j][k].ld;
        op.whole = true;
        op.term(F.H.p, F.H.ld, F.m, bs[j][k].p, bs[j][k].ld, ONE);
        auto& o = op.out(bk[j][k].p, bk[j][k].ld, F.m);
        out_term<T>(o, F.Ai.p, F.Ai.ld, F.m, bs[j][k].p, bs[j][k].ld, ONE);
        out_term<T>(o, F.W.p, F.W.ld, F.r, nullptr, 0, MONE);
        P.add(op);
      }
    }
    // every box's contribution to its parent's b
    P.stage();
    int nb = 1 << (2 * j);
    for (int nn = 0; nn < nb; ++nn) {
      FBox& Fn = B[j][nn];
      if (!Fn.r) continue;
      BoxOp<T> op;
      op.m = Fn.r;
      op.save = mr[j][nn].p;
      op.ldsave = mr[j][nn].ld;
      op.tile = tile_out();
      if (Fn.ill) {
        op.src = mr[j][nn].p;
        op.ldsrc = mr[j][nn].ld;
      } else if (Fn.m) {
        op.term(Fn.H.p, Fn.H.ld, Fn.m, bs[j][nn].p, bs[j][nn].ld, ONE);
      }
      for (int n = 0; n < 8; ++n) {
        int k = LS[nn].nb[n];
        if (k < 0 || B[j][k].color <= Fn.color || !B[j][k].m) continue;
        auto it = Fn.P2L.find(k);
        if (it == Fn.P2L.end() || !it->second.size()) continue;
        need_shape(it->second, Fn.r, B[j][k].m, "P2L");
        const Blk<double2>& v = B[j][k].ill ? bk[j][k] : bs[j][k];
        op.term(it->second.p, it->second.ld, it->second.c, v.p, v.ld, MONE);
      }
      P.add(op);
    }
  }
}

template <class T>
void Factor<T>::emit_base(Program<T>& P) {
  const int nbase = boff.empty() ? 0 : boff[16];
  if (nbase) {
    P.stage();
    BoxOp<T> op;
    op.m = 0;
    out_term<T>(op.out(xb.p, xb.ld, nbase), binv.p, binv.ld, nbase, yb.p, yb.ld, ONE);
    P.add(op);
  }
}

// back-substitution of levels jlo..jhi
template <class T>
void Factor<T>::emit_backward(Program<T>& P, int jlo, int jhi) {
  for (int j = jlo; j <= jhi; ++j) {
    const auto& LS = pb.lists[j];
    for (int c = 8; c >= 0; --c) {
      P.stage();
      for (int k : colors[j][c]) {
        FBox& F = B[j][k];
        if (!F.m) continue;
        BoxOp<T> op;
        op.m = F.m;
        op.src = bs[j][k].p;
        op.ldsrc = bs[j][k].ld;
        op.save = yt[j][k].p;
        op.ldsave = yt[j][k].ld;
        op.tile = tile_colour();
        for (int n = 0; n < 8; ++n) {
          int nn = LS[k].nb[n];
          if (nn < 0) continue;
          const FBox& Fn = B[j][nn];
          if (Fn.color > c) {
            if (!Fn.m) continue;
            auto it = F.P2P.find(nn);
            if (it == F.P2P.end() || !it->second.size()) continue;
            need_shape(it->second, F.m, Fn.m, "P2P (backward)");
            if (Fn.ill) {
              op.term(it->second.p, it->second.ld, it->second.c, xk[j][nn].p, xk[j][nn].ld, MONE);
            } else {
              op.term(it->second.p, it->second.ld, it->second.c, yt[j][nn].p, yt[j][nn].ld, MONE);
              auto iu = F.PU.find(nn);
              if (iu != F.PU.end() && iu->second.size()) {
                need_shape(iu->second, F.m, yk[j][nn].r, "P2P U (backward)");
                op.term(iu->second.p, iu->second.ld, iu->second.c, yk[j][nn].p, yk[j][nn].ld, ONE);
              }
            }
          } else {
            auto it = F.M2P.find(nn);
            if (it == F.M2P.end() || !it->second.size()) continue;
            need_shape(it->second, F.m, yk[j][nn].r, "M2P");
            op.term(it->second.p, it->second.ld, it->second.c, yk[j][nn].p, yk[j][nn].ld, MONE);
          }
        }
        P.add(op);
      }
      // ill boxes of this colour: v_k = yp_k - P2M A^{-1} y ; x_k = A^{-1} y - U v_k
      bool staged = false;
      for (int k : colors[j][c]) {
        FBox& F = B[j][k];
        if (!F.ill || !F.m) continue;
        if (!staged) P.stage(), staged = true;
        BoxOp<T> op;
        op.m = F.r;
        op.src = yk[j][k].p;
        op.ldsrc = yk[j][k].ld;
        op.save = vk[j][k].p;
        op.ldsave = vk[j][k].ld;
        op.whole = true;
        op.term(F.PA.p, F.PA.ld, F.m, yt[j][k].p, yt[j][k].ld, ONE);
        auto& o = op.out(xk[j][k].p, xk[j][k].ld, F.m);
        out_term<T>(o, F.Ai.p, F.Ai.ld, F.m, yt[j][k].p, yt[j][k].ld, ONE);
        out_term<T>(o, F.U.p, F.U.ld, F.r, nullptr, 0, MONE);
        P.add(op);
      }
    }
    // x of every regular box
    P.stage();
    for (int k = 0; k < (1 << (2 * j)); ++k) {
      FBox& F = B[j][k];
      if (!F.m || F.ill) continue;
      BoxOp<T> op;
      op.m = F.m;
      op.save = xk[j][k].p;
      op.ldsave = xk[j][k].ld;
      op.tile = tile_out();
      op.term(F.Z.p, F.Z.ld, F.m, yt[j][k].p, yt[j][k].ld, ONE);
      if (F.r) op.term(F.U.p, F.U.ld, F.r, yk[j][k].p, yk[j][k].ld, MONE);
      P.add(op);
    }
  }
}

// The top of the sweep (the forward sweep at levels jd..2, the base solve and
// the back-substitution at levels 2..jd) maps the stacked right-hand sides b_k
// of the level-jd boxes linearly to their solutions x_k and reads nothing
// else.  Its matrix D (n2 x n2, n2 = sum of the level-jd box sizes) is formed
// once per factorisation, by running that part of the sweep on the n2 columns
// of the identity; the solve sweeps then apply it as one dense product
// instead of about 38 short dependent launches per level.  jd is 3 when its
// n2 is at most kDenseTopMax, else 2 under the same bound, else the sweep
// stays staged.  dense_top_alloc sizes and allocates D (build_plan only needs
// where it lives); dense_top_fill forms it after build_plan, whose host work
// overlaps the sweep-operator kernels.  AIFMM_DENSE_TOP=0 keeps the staged
// form, AIFMM_DENSE_TOP=2 stops at level 2.
template <class T>
void Factor<T>::dense_top_alloc() {
  d2 = Blk<T>{};
  n2 = 0;
  st.dense_level = st.dense_n = 0;
  static const int top = [] {
    const char* e = getenv("AIFMM_DENSE_TOP");
    return e && e[0] == '0' ? 0 : e && e[0] == '2' ? 2 : 3;
  }();
  for (int j = top; j >= 2; --j) {
    if (L < j + 2) continue;
    const int nb = 1 << (2 * j);
    std::vector<int> off(nb + 1, 0);
    for (int k = 0; k < nb; ++k) off[k + 1] = off[k] + B[j][k].m;
    if (off[nb] == 0 || off[nb] > kDenseTopMax) continue;
    jd = j;
    off2 = off;
    n2 = off[nb];
    d2 = new_blk<T>(ar, n2, n2);
    st.dense_level = jd;
    st.dense_n = n2;
    return;
  }
}

template <class T>
void Factor<T>::dense_top_fill() {
  if (!n2) return;
  const int n = n2;
  Blk<double2> b2 = new_blk<double2>(sc, n, n), x2 = new_blk<double2>(sc, n, n);
  alloc_vectors(sc, n, jd, &b2, &x2);  // the sweep program already holds its own
  Program<T> P(n);
  emit_forward(P, 2, jd);
  emit_base(P);
  emit_backward(P, 2, jd);
  P.finalize(sc, pb.stream);
  // 256 threads, 2 columns per pass: the fastest launch shape for these runs
  // on MI355X (N = 147,456: level 3, 2,485 columns, 7.3 ms against 16.1 ms at
  // the default shape; level 2, 790 columns, 0.8 against 1.5 ms)
  P.force_shape(n, 256 + 2 * PROG_SHAPE_CH);
  CopyBatch<double2> cp;
  cp.add(b2, b2, 3);
  cp.run(sc, pb.stream);
  P.run_once(n, pb.stream);
  if constexpr (std::is_same<T, double2>::value) {
    cp.add(x2, d2);
    cp.run(sc, pb.stream);
  } else {
    CvtBatch<T> cv;
    cv.add(x2.p, d2.p, (int64_t)n * n);
    cv.run(sc, pb.stream);
  }
  HIP_OK(hipStreamSynchronize(pb.stream));  // before P and its launches go
}

template <class T>
void Factor<T>::build_plan(int cap) {
  plan.reset(new Program<T>(cap));
  pv.reset();
  plan_cap = cap;
  Program<T>& P = *plan;
  // The dense top serves sweeps of up to 64 columns (AIFMM_DENSE_COLS): at
  // N = 147,456, 16 columns took 22.6 ms per refined solve with it against
  // 23.9 ms staged, in the same 3 refinement steps; 256 columns gain nothing.
  static const int dense_cols = env_int("AIFMM_DENSE_COLS", 64);
  if (n2 && cap <= dense_cols) {
    Blk<double2> b2 = new_blk<double2>(pv, n2, cap), x2 = new_blk<double2>(pv, n2, cap);
    alloc_vectors(pv, cap, L, &b2, &x2);
    emit_forward(P, jd + 1, L);
    // levels jd..2 and the base solve: x2 = D b2, in row tiles
    P.stage();
    BoxOp<T> op;
    op.m = n2;
    op.save = x2.p;
    op.ldsave = x2.ld;
    op.term(d2.p, d2.ld, n2, b2.p, b2.ld, ONE);
    P.add(op);
    emit_backward(P, jd + 1, L);
  } else {
    alloc_vectors(pv, cap, L, nullptr, nullptr);
    emit_forward(P, 2, L);
    emit_base(P);
    emit_backward(P, 2, L);
  }
  P.finalize(pv, pb.stream);
}

template <class T>
void Factor<T>::swap_sweep() {
  pv.swap(spare_.pv);
  plan.swap(spare_.plan);
  std::swap(plan_cap, spare_.plan_cap);
  std::swap(Bt, spare_.Bt);
  std::swap(Xt, spare_.Xt);
  std::swap(yb, spare_.yb);
  std::swap(xb, spare_.xb);
  bk.swap(spare_.bk);
  bs.swap(spare_.bs);
  mr.swap(spare_.mr);
  xk.swap(spare_.xk);
  yk.swap(spare_.yk);
  yt.swap(spare_.yt);
  vk.swap(spare_.vk);
}

// One-column sweeps run the program built for one column; wider sweeps run
// one built for at least their width (rebuilt when a wider one comes).  The
// two programs and their vectors stay side by side.
template <class T>
Program<T>& Factor<T>::sweep_program(int nrhs) {
  if (B.empty()) throw std::runtime_error("solve before factor");
  nrhs = std::max(nrhs, 1);
  if ((nrhs == 1) != (plan_cap == 1) && plan) swap_sweep();
  if (nrhs == 1 ? plan_cap != 1 : nrhs > plan_cap) build_plan(nrhs);
  return *plan;
}

// b and Bt/Xt share the leading dimension N (complex128 in either precision)
template <class T>
void Factor<T>::sweep_load(const double2* b, int nrhs, hipStream_t s) {
  HIP_OK(hipMemcpyAsync(Bt.p, b, pb.N * nrhs * sizeof(double2), hipMemcpyDeviceToDevice, s));
}

template <class T>
void Factor<T>::sweep_store(double2* x, int nrhs, hipStream_t s) {
  HIP_OK(hipMemcpyAsync(x, Xt.p, pb.N * nrhs * sizeof(double2), hipMemcpyDeviceToDevice, s));
}

template <class T>
void Factor<T>::solve(const double2* b, double2* x, int nrhs) {
  Program<T>& P = sweep_program(nrhs);
  sweep_load(b, nrhs, pb.stream);
  P.run(nrhs, pb.stream, [&] { sweep_load(b, nrhs, pb.stream); });
  sweep_store(x, nrhs, pb.stream);
}

template <class T>
size_t Factor<T>::solve_bytes(int nrhs) const {
  if (plan) return (size_t)plan->bytes(nrhs);
  return 0;
}

template class Factor<double2>;
template class Factor<float2>;

}  // namespace aifmm
