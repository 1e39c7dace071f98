// FP64 hierarchical matvec: the nested H2 product of the operator built by
// getNodes / assemble_M2L / initialise_P2P_Leaf_Level
// (FMM2DTree_find_error.hpp:1105-1361, 1729-1741), as one box program:
//   upward,   level j = L..2:  mu_j[k] = P2M_k (x_k  or  stacked mu_{j+1} of the children)
//   downward, level j = 2..L:  lo_j[k] = sum_{b in IL(k)} M2L[k][b] mu_j[b]
//                                        + (L2P_parent lo_{j-1}[parent]) rows of k
//   leaves:                    y_k = L2P_k lo_L[k] + sum_{nn in N(k) and k} P2P[k][nn] x_nn
// (the near-field sum by k_near after the program, entries from the tables)
#include "h2.h"
#include "batch.h"
#include "kernels.h"

#include <vector>

namespace aifmm {

static const double2 ONE = {1.0, 0.0};
// offsets (dx, dy) of the near-field tables 0..7 (table 8: the box itself),
// the order of the neighbour slots and of the entry kernel's tables
static const int NEAR_OFF[8][2] = {{-1, -1}, {0, -1}, {1, -1}, {1, 0}, {1, 1}, {0, 1}, {-1, 1}, {-1, 0}};

void H2Op::build(int nrhs) {
  Problem& P = pb;
  const int L = P.L, R = P.R;
  const int N = (int)P.N;
  ws.reset();
  plan.reset(new Program<double2>(nrhs));
  cap = nrhs;
  xin = ws.get<double2>((size_t)N * nrhs);
  yout = ws.get<double2>((size_t)N * nrhs);
  const double2* x = xin;
  double2* y = yout;
  if (R > 64) throw std::runtime_error("near field kernel takes at most 64 points per leaf");
  std::vector<int> nbr((size_t)9 << (2 * L), -1);
  std::vector<std::vector<int>> off(L + 1);
  std::vector<Blk<double2>> Mu(L + 1), Lo(L + 1);
  for (int j = 2; j <= L; ++j) {
    int nb = 1 << (2 * j);
    off[j].assign(nb + 1, 0);
    for (int k = 0; k < nb; ++k) off[j][k + 1] = off[j][k] + P.sk[j][k].r;
    Mu[j] = new_blk<double2>(ws, off[j][nb], nrhs);
    Lo[j] = new_blk<double2>(ws, off[j][nb], nrhs);
  }
  Program<double2>& G = *plan;
  auto need = [](bool ok, const char* what) {
    if (!ok) throw std::runtime_error(std::string("H2 operator block shape mismatch: ") + what);
  };
  // upward pass
  for (int j = L; j >= 2; --j) {
    G.stage();
    int nb = 1 << (2 * j);
    for (int k = 0; k < nb; ++k) {
      const SkelBox& B = P.sk[j][k];
      if (!B.r) continue;
      BoxOp<double2> op;
      op.m = B.r;
      op.save = Mu[j].p + off[j][k];
      op.ldsave = Mu[j].ld;
      if (B.m) {
        need(B.P2M.r == B.r && B.P2M.c == B.m, "P2M");
        if (j == L) op.term(B.P2M.p, B.P2M.ld, B.m, x + (size_t)k * R, N, ONE);
        else op.term(B.P2M.p, B.P2M.ld, B.m, Mu[j + 1].p + off[j + 1][4 * k], Mu[j + 1].ld, ONE);
      }
      G.add(op);
    }
  }
  // downward pass with M2L; near field and L2P at the leaves
  for (int j = 2; j <= L; ++j) {
    G.stage();
    int nb = 1 << (2 * j);
    for (int k = 0; k < nb; ++k) {
      const SkelBox& B = P.sk[j][k];
      if (!B.r && j < L) continue;
      BoxOp<double2> op;
      op.m = B.r;
      if (B.r) {
        // the local vector is kept for the children (and is the tiled phase-1 target)
        op.save = Lo[j].p + off[j][k];
        op.ldsave = Lo[j].ld;
        const Lists& Ls = P.lists[j][k];
        for (int s = 0; s < 40; ++s) {
          int b = s < 16 ? Ls.inner[s] : Ls.outer[s - 16];
          if (b < 0) continue;
          auto it = B.M2L.find(b);
          if (it == B.M2L.end() || !it->second.size()) continue;
          need(it->second.r == B.r && it->second.c == P.sk[j][b].r, "M2L");
          op.term(it->second.p, it->second.ld, it->second.c, Mu[j].p + off[j][b], Mu[j].ld, ONE);
        }
        if (j > 2) {
          const SkelBox& Pa = P.sk[j - 1][k >> 2];
          int o = 0;
          for (int c = 0; c < (k & 3); ++c) o += P.sk[j][(k & ~3) + c].r;
          need(Pa.L2P.r == Pa.m && o + B.r <= Pa.L2P.r, "parent L2P");
          if (Pa.r) op.term(Pa.L2P.p + o, Pa.L2P.ld, Pa.r, Lo[j - 1].p + off[j - 1][k >> 2], Lo[j - 1].ld, ONE);
        }
      }
      if (j == L) {
        auto& out = op.out(y + (size_t)k * R, N, R);
        if (B.r) {
          need(B.L2P.r == R && B.L2P.c == B.r, "leaf L2P");
          out_term<double2>(out, B.L2P.p, B.L2P.ld, B.r, nullptr, 0, ONE);
        }
        // the near field (P2P) is added by k_near after the program, from
        // the tables: the same entries without reading the assembled blocks
        for (int slot = 0; slot < 9; ++slot) {
          int nn = slot == 8 ? k : P.lists[L][k].nb[slot];
          nbr[(size_t)k * 9 + slot] = -1;
          if (nn < 0) continue;
          auto it = P.P2P[k].find(nn);
          if (it == P.P2P[k].end() || !it->second.size()) continue;
          need(it->second.r == R && it->second.c == R, "P2P");
          // k_near takes table `slot` for this neighbour: its offset must be
          // the table's (the neighbour lists and the tables share the order)
          if (slot < 8) {
            int xk, yk, xn, yn;
            box_xy(L, k, xk, yk);
            box_xy(L, nn, xn, yn);
            if (xn - xk != NEAR_OFF[slot][0] || yn - yk != NEAR_OFF[slot][1])
              throw std::runtime_error("neighbour slot does not match its near-field table");
          }
          nbr[(size_t)k * 9 + slot] = nn;
        }
      }
      G.add(op);
    }
  }
  G.finalize(ws, P.stream);
  // kappa^2 T, each entry scaled as the assembled entries are (T kappa kappa)
  {
    const double kap = P.ctx.kappa;
    std::vector<double2> tk(P.T_h.size());
    for (size_t e = 0; e < tk.size(); ++e) {
      double2 v = P.T_h[e];
      v.x *= kap, v.y *= kap;
      v.x *= kap, v.y *= kap;
      tk[e] = v;
    }
    dtk = ws.get<double2>(tk.size());
    h2d(dtk, tk.data(), tk.size() * sizeof(double2), P.stream);
  }
  dnbr = ws.get<int>(nbr.size());
  h2d(dnbr, nbr.data(), nbr.size() * sizeof(int), P.stream);
}

NearDesc H2Op::near_desc(int nrhs) const {
  NearDesc d{};
  d.T = dtk;
  d.q = pb.ctx.q;
  d.nbr = dnbr;
  d.x = xin;
  d.y = yout;
  d.kappa = pb.ctx.kappa;
  d.L = pb.L;
  d.R = pb.R;
  d.nbox = 1 << (2 * pb.L);
  d.nrhs = nrhs;
  d.N = pb.N;
  return d;
}

void H2Op::near(int nrhs, hipStream_t s) {
  const NearDesc d = near_desc(nrhs);
  NearDesc* dd = (NearDesc*)stage_dev.alloc(sizeof(NearDesc));
  h2d(dd, &d, sizeof(d), s);
  launch_near(1, d.nbox, dd, s, d.nrhs);
  stage_dev.reset();
}

Program<double2>& H2Op::program(int nrhs) {
  if (!pb.assembled) pb.assemble();
  if (!plan || nrhs > cap) build(nrhs);
  return *plan;
}

void H2Op::load(const double2* x, int nrhs, hipStream_t s) {
  HIP_OK(hipMemcpyAsync(xin, x, (size_t)pb.N * nrhs * sizeof(double2), hipMemcpyDeviceToDevice, s));
}

void H2Op::store(double2* y, int nrhs, hipStream_t s) {
  HIP_OK(hipMemcpyAsync(y, yout, (size_t)pb.N * nrhs * sizeof(double2), hipMemcpyDeviceToDevice, s));
}

void H2Op::apply(const double2* x, double2* y, int nrhs) {
  Program<double2>& P = program(nrhs);
  load(x, nrhs, pb.stream);
  P.run(nrhs, pb.stream, [&] { load(x, nrhs, pb.stream); });
  near(nrhs, pb.stream);
  store(y, nrhs, pb.stream);
}

size_t H2Op::bytes(int nrhs) const {
  if (plan) return (size_t)plan->bytes(nrhs);
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
