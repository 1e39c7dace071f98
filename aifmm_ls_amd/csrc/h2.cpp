// FP64 hierarchical matvec: leaf near field, upward pass (P2M), M2L on every
// level, downward pass (L2P) — the nested H2 product of the operator built
// by getNodes / assemble_M2L / initialise_P2P_Leaf_Level
// (FMM2DTree_find_error.hpp:1105-1361, 1729-1741).
#include "h2.h"

#include <vector>

namespace aifmm {

static const double2 ONE = {1.0, 0.0}, ZERO = {0.0, 0.0};

void col_norms(Arena& sc, const double2* v, int64_t n, int nrhs, double* out, hipStream_t s) {
  std::vector<CopyDesc<double2>> d(nrhs);
  for (int c = 0; c < nrhs; ++c) {
    d[c] = CopyDesc<double2>{};
    d[c].src = v + (size_t)n * c;
    d[c].nr = (int)n;
    d[c].nc = 1;
    d[c].lds = (int)n;
  }
  auto* dd = upload(sc, d, s);
  double* dn = sc.get<double>(nrhs);
  launch_fro<double2>(nrhs, dd, dn, s);
  HIP_OK(hipMemcpyAsync(out, dn, 8 * nrhs, hipMemcpyDeviceToHost, s));
  HIP_OK(hipStreamSynchronize(s));
  sc.reset();
}

void H2Op::apply(const double2* x, double2* y, int nrhs) {
  Problem& P = pb;
  if (!P.assembled) P.assemble();
  hipStream_t s = P.stream;
  const int L = P.L, R = P.R;
  const int N = (int)P.N;
  ws.reset();
  // per-level multipole / local buffers, boxes stacked in box order
  std::vector<std::vector<int>> off(L + 1);
  std::vector<Blk<double2>> Mu(L + 1), Lo(L + 1);
  for (int j = 2; j <= L; ++j) {
    int nb = 1 << (2 * j);
    off[j].assign(nb + 1, 0);
    for (int k = 0; k < nb; ++k) off[j][k + 1] = off[j][k] + P.sk[j][k].r;
    Mu[j] = new_blk<double2>(ws, off[j][nb], nrhs);
    Lo[j] = new_blk<double2>(ws, off[j][nb], nrhs);
  }
  Blk<double2> X{(double2*)x, N, nrhs, N}, Y{y, N, nrhs, N};
  GemmBatch<double2> g;
  CopyBatch<double2> cp;
  // y = 0 ; locals = 0
  cp.add(Y, Y, 4);
  for (int j = 2; j <= L; ++j)
    if (Lo[j].size()) cp.add(Lo[j], Lo[j], 4);
  cp.run(sc, s);
  // near field, one batch per neighbour slot (outputs written once per batch)
  for (int slot = 0; slot < 9; ++slot) {
    for (int k = 0; k < P.nleaf; ++k) {
      int nn = slot == 8 ? k : P.lists[L][k].nb[slot];
      if (nn < 0) continue;
      const Blk<double2>& A = P.P2P[k].at(nn);
      g.add(A, 0, X.sub(nn * R, 0, R, nrhs), 0, Y.sub(k * R, 0, R, nrhs), ONE, ONE);
    }
    g.run(sc, s);
  }
  // upward pass
  for (int j = L; j >= 2; --j) {
    int nb = 1 << (2 * j);
    for (int k = 0; k < nb; ++k) {
      const SkelBox& B = P.sk[j][k];
      if (!B.r || !B.m) continue;
      Blk<double2> src = (j == L) ? X.sub(k * R, 0, R, nrhs) : Mu[j + 1].sub(off[j + 1][4 * k], 0, B.m, nrhs);
      g.add(B.P2M, 0, src, 0, Mu[j].sub(off[j][k], 0, B.r, nrhs), ONE, ZERO);
    }
    g.run(sc, s);
  }
  // M2L on each level, one batch per interaction-list slot
  for (int j = 2; j <= L; ++j) {
    int nb = 1 << (2 * j);
    for (int slot = 0; slot < 40; ++slot) {
      for (int k = 0; k < nb; ++k) {
        const Lists& Ls = P.lists[j][k];
        int b = slot < 16 ? Ls.inner[slot] : Ls.outer[slot - 16];
        if (b < 0) continue;
        const SkelBox& B = P.sk[j][k];
        const Blk<double2>& M = B.M2L.at(b);
        if (!M.size()) continue;
        g.add(M, 0, Mu[j].sub(off[j][b], 0, P.sk[j][b].r, nrhs), 0, Lo[j].sub(off[j][k], 0, B.r, nrhs), ONE, ONE);
      }
      g.run(sc, s);
    }
  }
  // downward pass: children locals += L2P_parent * local_parent ; leaf output
  for (int j = 2; j <= L; ++j) {
    int nb = 1 << (2 * j);
    for (int k = 0; k < nb; ++k) {
      const SkelBox& B = P.sk[j][k];
      if (!B.r || !B.m) continue;
      Blk<double2> dst = (j == L) ? Y.sub(k * R, 0, R, nrhs) : Lo[j + 1].sub(off[j + 1][4 * k], 0, B.m, nrhs);
      g.add(B.L2P, 0, Lo[j].sub(off[j][k], 0, B.r, nrhs), 0, dst, ONE, ONE);
    }
    g.run(sc, s);
  }
  HIP_OK(hipStreamSynchronize(s));
  sc.reset();
}

size_t H2Op::bytes(int nrhs) const {
  size_t e = 0;
  for (auto& m : pb.P2P)
    for (auto& b : m) e += b.second.size();
  for (int j = 2; j <= pb.L; ++j)
    for (auto& B : pb.sk[j]) {
      e += B.L2P.size() + B.P2M.size();
      for (auto& m : B.M2L) e += m.second.size();
    }
  return e * 16 + (size_t)pb.N * nrhs * 32;
}

}  // namespace aifmm
