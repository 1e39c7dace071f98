#include "krylov.h"

#include "batch.h"

#include <algorithm>
#include <cmath>
#include <complex>

namespace aifmm {

using cplx = std::complex<double>;

void Refiner::norms(const double2* v, int nrhs, std::vector<double>& out) {
  const int64_t N = pb_.N;
  const int nblk = resid_blocks(N);
  std::vector<double> part((size_t)nblk * nrhs);
  double* dp = (double*)dpart_;
  launch_resid(v, nullptr, nullptr, N, nrhs, dp, pb_.stream);
  d2h(part.data(), dp, part.size() * 8, pb_.stream);
  HIP_OK(hipStreamSynchronize(pb_.stream));
  out.assign(nrhs, 0);
  for (int c = 0; c < nrhs; ++c) {
    double s = 0;
    for (int b = 0; b < nblk; ++b) s += part[(size_t)c * nblk + b];
    out[c] = std::sqrt(s);
  }
}

void Refiner::dots(const double2* V, int64_t vstride, int nv, const double2* w, int nrhs, std::vector<double2>& out) {
  const int64_t N = pb_.N;
  const int nblk = resid_blocks(N);
  std::vector<double2> part((size_t)nblk * nv * nrhs);
  launch_multidot(V, vstride, nv, w, N, nrhs, dpart_, pb_.stream);
  d2h(part.data(), dpart_, part.size() * 16, pb_.stream);
  HIP_OK(hipStreamSynchronize(pb_.stream));
  out.assign((size_t)nv * nrhs, make_double2(0, 0));
  for (int c = 0; c < nrhs; ++c)
    for (int i = 0; i < nv; ++i) {
      double2 s = make_double2(0, 0);
      for (int b = 0; b < nblk; ++b) {
        const double2 p = part[((size_t)c * nv + i) * nblk + b];
        s.x += p.x;
        s.y += p.y;
      }
      out[(size_t)c * nv + i] = s;
    }
}

void Refiner::lincomb(double2* w, const double2* V, int64_t vstride, int nv, const std::vector<double2>& coef,
                      const std::vector<double2>& beta, int nrhs) {
  // coefficients go through the pinned staging ring: the copies are queued
  // without waiting, and the host vectors may go out of scope on return
  if (nv) {
    const size_t b = (size_t)nv * nrhs * 16;
    h2d(dcoef_, coef.data(), b, pb_.stream);
  }
  const size_t bb = (size_t)nrhs * 16;
  h2d(dbeta_, beta.data(), bb, pb_.stream);
  launch_lincomb(w, V, vstride, nv, dcoef_, dbeta_, pb_.N, nrhs, pb_.stream);
}

RefineResult Refiner::run(const Precond& M, const double2* b, double2* x, int nr, double tol, int maxit, int method,
                          bool x_given) {
  const int64_t N = pb_.N;
  hipStream_t s = pb_.stream;
  const size_t vb = (size_t)N * nr;
  RefineResult R;
  if (!x_given) M(b, x);
  if (maxit <= 0) return R;
  // restart length: up to 20, the Krylov vectors (V and Z) within 32 GiB of
  // the 288 GB HBM (at 256 right-hand sides of N = 589,824: 5)
  int m = std::min(maxit, 20);
  const size_t lim = (size_t(32) << 30) / (vb * 32);
  m = std::max(1, std::min<int>(m, (int)std::max<size_t>(lim, 2) - 1));
  ws_.reset();
  double2* V = ws_.get<double2>((size_t)(m + 1) * vb);
  double2* Z = ws_.get<double2>((size_t)m * vb);
  double2* r = ws_.get<double2>(vb);
  dpart_ = ws_.get<double2>((size_t)resid_blocks(N) * (m + 1) * nr);
  dcoef_ = ws_.get<double2>((size_t)(m + 1) * nr);
  dbeta_ = ws_.get<double2>(nr);

  std::vector<double> bn, rn;
  norms(b, nr, bn);
  auto relmax = [&](const std::vector<double>& v) {
    double q = 0;
    for (int c = 0; c < nr; ++c) q = std::max(q, bn[c] > 0 ? v[c] / bn[c] : v[c]);
    return q;
  };
  const double2 one = make_double2(1, 0);
  for (;;) {
    // true residual r = b - A x
    A_.apply(x, r, nr);
    {
      const int nblk = resid_blocks(N);
      std::vector<double> part((size_t)nblk * nr);
      launch_resid(b, r, r, N, nr, (double*)dpart_, s);
      d2h(part.data(), dpart_, part.size() * 8, s);
      HIP_OK(hipStreamSynchronize(s));
      rn.assign(nr, 0);
      for (int c = 0; c < nr; ++c) {
        double q = 0;
        for (int k = 0; k < nblk; ++k) q += part[(size_t)c * nblk + k];
        rn[c] = std::sqrt(q);
      }
    }
    R.relres = relmax(rn);
    if (R.relres <= tol || R.steps >= maxit) break;
    if (method == 1) {  // Richardson: x += M r
      M(r, Z);
      launch_axpy(x, Z, (int64_t)vb, s);
      ++R.steps;
      continue;
    }
    // FGMRES cycle
    HIP_OK(hipMemcpyAsync(V, r, vb * 16, hipMemcpyDeviceToDevice, s));
    std::vector<double2> sc(nr);
    for (int c = 0; c < nr; ++c) sc[c] = make_double2(rn[c] > 0 ? 1.0 / rn[c] : 0.0, 0);
    lincomb(V, nullptr, 0, 0, {}, sc, nr);
    std::vector<cplx> H((size_t)(m + 1) * m * nr, 0.0), g((size_t)(m + 1) * nr, 0.0), cs((size_t)m * nr, 0.0),
        sn((size_t)m * nr, 0.0);
    auto Hc = [&](int c, int i, int j) -> cplx& { return H[((size_t)c * (m + 1) + i) * m + j]; };
    for (int c = 0; c < nr; ++c) g[(size_t)c * (m + 1)] = rn[c];
    int k = 0;
    double est_end = -1;  // the recurrence's residual when the cycle converged
    for (int j = 0; j < m && R.steps < maxit; ++j) {
      double2* vj = V + (size_t)j * vb;
      double2* zj = Z + (size_t)j * vb;
      double2* w = V + (size_t)(j + 1) * vb;
      M(vj, zj);
      A_.apply(zj, w, nr);
      ++R.steps;
      std::vector<double2> h, h2w, neg((size_t)(j + 1) * nr);
      std::vector<double2> keep(nr, one);
      dots(V, (int64_t)vb, j + 1, w, nr, h);
      for (size_t q = 0; q < neg.size(); ++q) neg[q] = make_double2(-h[q].x, -h[q].y);
      lincomb(w, V, (int64_t)vb, j + 1, neg, keep, nr);
      // second pass: the products with V_0..V_j and, as basis slot j + 1 is w
      // itself, |w|^2; then |w - V h2|^2 = |w|^2 - |h2|^2 (V orthonormal)
      // gives the new norm, and one update both projects and normalises
      dots(V, (int64_t)vb, j + 2, w, nr, h2w);
      std::vector<double2> h2((size_t)(j + 1) * nr);
      std::vector<double> hn(nr);
      bool exact = false;
      for (int c = 0; c < nr; ++c) {
        double w2 = h2w[(size_t)c * (j + 2) + j + 1].x, p2 = 0;
        for (int i = 0; i <= j; ++i) {
          const double2 v = h2w[(size_t)c * (j + 2) + i];
          h2[(size_t)c * (j + 1) + i] = v;
          p2 += v.x * v.x + v.y * v.y;
        }
        const double r2 = w2 - p2;
        if (!(r2 > 1e-8 * w2)) exact = true;  // cancellation: measure instead
        hn[c] = std::sqrt(std::max(r2, 0.0));
      }
      for (size_t q = 0; q < neg.size(); ++q) neg[q] = make_double2(-h2[q].x, -h2[q].y);
      if (exact) {
        lincomb(w, V, (int64_t)vb, j + 1, neg, keep, nr);
        norms(w, nr, hn);
        for (int c = 0; c < nr; ++c) sc[c] = make_double2(hn[c] > 0 ? 1.0 / hn[c] : 0.0, 0);
        lincomb(w, nullptr, 0, 0, {}, sc, nr);
      } else {
        for (int c = 0; c < nr; ++c) {
          const double s = hn[c] > 0 ? 1.0 / hn[c] : 0.0;
          sc[c] = make_double2(s, 0);
          for (int i = 0; i <= j; ++i) {
            double2& v = neg[(size_t)c * (j + 1) + i];
            v.x *= s;
            v.y *= s;
          }
        }
        lincomb(w, V, (int64_t)vb, j + 1, neg, sc, nr);
      }
      double worst = 0;
      for (int c = 0; c < nr; ++c) {
        for (int i = 0; i <= j; ++i) {
          const double2 a = h[(size_t)c * (j + 1) + i], e = h2[(size_t)c * (j + 1) + i];
          Hc(c, i, j) = cplx(a.x + e.x, a.y + e.y);
        }
        Hc(c, j + 1, j) = hn[c];
        // previous rotations
        for (int i = 0; i < j; ++i) {
          const cplx ci = cs[(size_t)c * m + i], si = sn[(size_t)c * m + i];
          const cplx a = Hc(c, i, j), bb = Hc(c, i + 1, j);
          Hc(c, i, j) = ci * a + si * bb;
          Hc(c, i + 1, j) = -std::conj(si) * a + ci * bb;
        }
        // new rotation annihilating H(j+1, j)
        const cplx h1 = Hc(c, j, j), h2v = Hc(c, j + 1, j);
        const double t = std::sqrt(std::norm(h1) + std::norm(h2v));
        cplx cj = 0, sj = 1;
        if (t > 0 && std::abs(h1) > 0) {
          cj = std::abs(h1) / t;
          sj = (h1 / std::abs(h1)) * std::conj(h2v) / t;
        }
        cs[(size_t)c * m + j] = cj;
        sn[(size_t)c * m + j] = sj;
        Hc(c, j, j) = cj * h1 + sj * h2v;
        Hc(c, j + 1, j) = 0;
        cplx& gj = g[(size_t)c * (m + 1) + j];
        g[(size_t)c * (m + 1) + j + 1] = -std::conj(sj) * gj;
        gj = cj * gj;
        const double est = std::abs(g[(size_t)c * (m + 1) + j + 1]);
        worst = std::max(worst, bn[c] > 0 ? est / bn[c] : est);
      }
      k = j + 1;
      // Stop bound: with the recurrence trusted (tol >= 1e-11, below) it
      // only needs a margin for FP64 rounding; a tighter tol is confirmed by
      // a true residual, and half of it leaves room for the rounding floor.
      if (worst <= (tol >= 1e-11 ? 0.9 : 0.5) * tol) {
        est_end = worst;
        break;
      }
    }
    // y = H[:k, :k]^{-1} g[:k] per column; x += Z y
    std::vector<double2> y((size_t)k * nr);
    for (int c = 0; c < nr; ++c) {
      std::vector<cplx> yy(k, 0.0);
      for (int i = k - 1; i >= 0; --i) {
        cplx acc = g[(size_t)c * (m + 1) + i];
        for (int l = i + 1; l < k; ++l) acc -= Hc(c, i, l) * yy[l];
        const cplx d = Hc(c, i, i);
        yy[i] = std::abs(d) > 0 ? acc / d : 0.0;
      }
      for (int i = 0; i < k; ++i) y[(size_t)c * k + i] = make_double2(yy[i].real(), yy[i].imag());
    }
    lincomb(x, Z, (int64_t)vb, k, y, std::vector<double2>(nr, one), nr);
    // Converged by the recurrence: its residual is |b - A x| of the FP64
    // operator up to FP64 rounding (w = A z is formed explicitly and V is
    // kept orthonormal by two Gram-Schmidt passes), so the 10 % margin needs
    // no further product to confirm while tol is far above the rounding
    // floor (about 1e-15; the recurrence keeps shrinking below it).
    // A cycle that ended on its restart length or the step limit, or a
    // tighter tol, measures the true residual above.
    if (est_end >= 0 && tol >= 1e-11) {
      R.relres = est_end;
      break;
    }
  }
  HIP_OK(hipStreamSynchronize(s));
  return R;
}

std::vector<RefineResult> BatchRefiner::run(const std::vector<BatchVec>& v, const Apply& M, const Apply& A, double tol,
                                            int maxit, hipStream_t s) {
  const int n = (int)v.size();
  std::vector<RefineResult> R(n);
  std::vector<const double2*> src(n);
  std::vector<double2*> dst(n);
  for (int i = 0; i < n; ++i) src[i] = v[i].b, dst[i] = v[i].x;
  M(src, dst);  // x = M b
  if (maxit <= 0 || n == 0) return R;
  int m = std::min(maxit, 20);
  ws_.reset();
  struct Mem {
    double2 *V, *Z, *r;
    int nblk;
    size_t poff;  // offset of this member's partials
  };
  std::vector<Mem> mem(n);
  size_t ptot = 0;
  for (int i = 0; i < n; ++i) {
    const size_t N = (size_t)v[i].N;
    mem[i].V = ws_.get<double2>((m + 1) * N);
    mem[i].Z = ws_.get<double2>(m * N);
    mem[i].r = ws_.get<double2>(N);
    mem[i].nblk = resid_blocks(v[i].N);
    mem[i].poff = ptot;
    ptot += (size_t)mem[i].nblk * (m + 2);
  }
  double2* dpart = ws_.get<double2>(ptot);
  double2* dcoef = ws_.get<double2>((size_t)n * (m + 1));
  double2* dbeta = ws_.get<double2>(n);
  double2* dsum = ws_.get<double2>((size_t)n * (m + 2));
  int64_t* doff = ws_.get<int64_t>(n);
  int* dnb = ws_.get<int>(n);
  {
    std::vector<int64_t> off(n);
    std::vector<int> nb(n);
    for (int i = 0; i < n; ++i) off[i] = (int64_t)mem[i].poff, nb[i] = mem[i].nblk;
    h2d(doff, off.data(), n * 8, s);
    h2d(dnb, nb.data(), n * 4, s);
  }
  // member descriptors of the batched vector launches (one launch per
  // operation for all members; uploads and launches share the stream, so the
  // one device array is rewritten only after the previous launch read it)
  VecMember* dmem = ws_.get<VecMember>(n);
  int maxblk = 1;
  for (int i = 0; i < n; ++i) maxblk = std::max(maxblk, mem[i].nblk);
  std::vector<VecMember> vm(n);
  auto send = [&]() {
    const size_t bytes = (size_t)n * sizeof(VecMember);
    h2d(dmem, vm.data(), bytes, s);
  };
  // partial squared norms of b_i - ax_i (ax_i null: of b_i; r_i null: not stored)
  auto resid = [&](const std::vector<const double2*>& b, const std::vector<const double2*>& ax,
                   const std::vector<double2*>& r) {
    for (int i = 0; i < n; ++i) {
      vm[i] = VecMember{};
      vm[i].b = b[i];
      vm[i].ax = ax.empty() ? nullptr : ax[i];
      vm[i].r = r.empty() ? nullptr : r[i];
      vm[i].n = v[i].N;
      vm[i].part = dpart + mem[i].poff;
    }
    send();
    launch_resid_batch(n, maxblk, dmem, s);
  };
  // partial dots of w_i against V_i[0..nv)
  auto multidot = [&](const std::vector<double2*>& w, int nv) {
    for (int i = 0; i < n; ++i) {
      vm[i] = VecMember{};
      vm[i].V = mem[i].V;
      vm[i].vstride = v[i].N;
      vm[i].w = w[i];
      vm[i].n = v[i].N;
      vm[i].part = dpart + mem[i].poff;
      vm[i].nv = nv;
    }
    send();
    launch_multidot_batch(n, maxblk, dmem, s);
  };
  // w_i = dbeta[i] w_i + sum_q dcoef[i (m + 1) + q] B_i[q] (B_i = V_i or Z_i)
  auto lincomb_dev = [&](const std::vector<double2*>& w, const std::vector<int>& nv, bool from_z) {
    for (int i = 0; i < n; ++i) {
      vm[i] = VecMember{};
      vm[i].V = nv[i] ? (from_z ? mem[i].Z : mem[i].V) : nullptr;
      vm[i].vstride = v[i].N;
      vm[i].w = w[i];
      vm[i].n = v[i].N;
      vm[i].coef = dcoef + (size_t)i * (m + 1);
      vm[i].beta = dbeta + i;
      vm[i].nv = nv[i];
    }
    send();
    launch_lincomb_batch(n, maxblk, dmem, s);
  };
  std::vector<double2> sums((size_t)n * (m + 2));
  // the members' partial sums reduced on the device, then nv values each to
  // the host: out[i][0..nv)
  auto fetch = [&](int nv, std::vector<std::vector<double2>>& out) {
    launch_sum_partials(dpart, doff, dnb, n, nv, 0, dsum, s);
    d2h(sums.data(), dsum, (size_t)n * nv * 16, s);
    HIP_OK(hipStreamSynchronize(s));
    out.assign(n, {});
    for (int i = 0; i < n; ++i) out[i].assign(sums.begin() + (size_t)i * nv, sums.begin() + (size_t)(i + 1) * nv);
  };
  // 2-norms from the real partials of launch_resid
  auto fetch_norms = [&](std::vector<double>& out) {
    launch_sum_partials(dpart, doff, dnb, n, 1, 1, dsum, s);
    d2h(sums.data(), dsum, (size_t)n * 16, s);
    HIP_OK(hipStreamSynchronize(s));
    out.resize(n);
    for (int i = 0; i < n; ++i) out[i] = std::sqrt(sums[i].x);
  };
  // w_i = beta_i w_i + sum_q coef_i[q] V_i[q] for every member (coefficients
  // staged through the pinned ring; no host wait)
  auto lincomb = [&](const std::vector<double2*>& w, const std::vector<int>& nv,
                     const std::vector<std::vector<double2>>& coef, const std::vector<double2>& beta) {
    std::vector<double2> hc((size_t)n * (m + 1), make_double2(0, 0));
    for (int i = 0; i < n; ++i)
      for (int q = 0; q < nv[i]; ++q) hc[(size_t)i * (m + 1) + q] = coef[i][q];
    h2d(dcoef, hc.data(), hc.size() * 16, s);
    h2d(dbeta, beta.data(), n * 16, s);
    lincomb_dev(w, nv, false);
  };
  std::vector<double> bn(n), rn(n);
  std::vector<const double2*> bs(n);
  for (int i = 0; i < n; ++i) bs[i] = v[i].b;
  resid(bs, {}, {});
  fetch_norms(bn);
  int steps = 0;
  for (;;) {
    // true residuals r = b - A x
    for (int i = 0; i < n; ++i) src[i] = v[i].x, dst[i] = mem[i].r;
    A(src, dst);
    {
      std::vector<const double2*> ax(n);
      std::vector<double2*> rs(n);
      for (int i = 0; i < n; ++i) ax[i] = rs[i] = mem[i].r;
      resid(bs, ax, rs);
    }
    fetch_norms(rn);
    double worst = 0;
    for (int i = 0; i < n; ++i) {
      R[i].relres = bn[i] > 0 ? rn[i] / bn[i] : rn[i];
      R[i].steps = steps;
      worst = std::max(worst, R[i].relres);
    }
    if (worst <= tol || steps >= maxit) break;
    // V_0 = r / |r|
    std::vector<int> nv0(n, 0);
    std::vector<std::vector<double2>> none(n);
    std::vector<double2> sc(n);
    for (int i = 0; i < n; ++i) {
      HIP_OK(hipMemcpyAsync(mem[i].V, mem[i].r, v[i].N * 16, hipMemcpyDeviceToDevice, s));
      sc[i] = make_double2(rn[i] > 0 ? 1.0 / rn[i] : 0.0, 0);
      dst[i] = mem[i].V;
    }
    lincomb(dst, nv0, none, sc);
    std::vector<std::vector<std::complex<double>>> H(n, std::vector<std::complex<double>>((size_t)(m + 1) * m, 0.0)),
        g(n, std::vector<std::complex<double>>(m + 1, 0.0)), cs(n, std::vector<std::complex<double>>(m, 0.0)),
        sn(n, std::vector<std::complex<double>>(m, 0.0));
    for (int i = 0; i < n; ++i) g[i][0] = rn[i];
    int k = 0;
    for (int j = 0; j < m && steps < maxit; ++j) {
      for (int i = 0; i < n; ++i) src[i] = mem[i].V + (size_t)j * v[i].N, dst[i] = mem[i].Z + (size_t)j * v[i].N;
      M(src, dst);
      for (int i = 0; i < n; ++i) src[i] = mem[i].Z + (size_t)j * v[i].N, dst[i] = mem[i].V + (size_t)(j + 1) * v[i].N;
      A(src, dst);
      ++steps;
      std::vector<int> nvj(n, j + 1);
      std::vector<double2*> w(n);
      for (int i = 0; i < n; ++i) w[i] = mem[i].V + (size_t)(j + 1) * v[i].N;
      multidot(w, j + 1);
      std::vector<std::vector<double2>> h, h2w, neg(n);
      fetch(j + 1, h);
      std::vector<double2> one(n, make_double2(1, 0));
      for (int i = 0; i < n; ++i) {
        neg[i].resize(j + 1);
        for (int q = 0; q <= j; ++q) neg[i][q] = make_double2(-h[i][q].x, -h[i][q].y);
      }
      lincomb(w, nvj, neg, one);
      // second pass with |w|^2 (w is basis slot j + 1)
      multidot(w, j + 2);
      fetch(j + 2, h2w);
      std::vector<double> hn(n);
      bool exact = false;
      for (int i = 0; i < n; ++i) {
        double p2 = 0;
        for (int q = 0; q <= j; ++q) p2 += h2w[i][q].x * h2w[i][q].x + h2w[i][q].y * h2w[i][q].y;
        const double w2 = h2w[i][j + 1].x, r2 = w2 - p2;
        if (!(r2 > 1e-8 * w2)) exact = true;
        hn[i] = std::sqrt(std::max(r2, 0.0));
        for (int q = 0; q <= j; ++q) neg[i][q] = make_double2(-h2w[i][q].x, -h2w[i][q].y);
      }
      if (exact) {
        lincomb(w, nvj, neg, one);
        resid(std::vector<const double2*>(w.begin(), w.end()), {}, {});
        fetch_norms(hn);
        for (int i = 0; i < n; ++i) sc[i] = make_double2(hn[i] > 0 ? 1.0 / hn[i] : 0.0, 0);
        lincomb(w, nv0, none, sc);
      } else {
        for (int i = 0; i < n; ++i) {
          const double sI = hn[i] > 0 ? 1.0 / hn[i] : 0.0;
          sc[i] = make_double2(sI, 0);
          for (auto& c : neg[i]) c.x *= sI, c.y *= sI;
        }
        lincomb(w, nvj, neg, sc);
      }
      double worst_est = 0;
      for (int i = 0; i < n; ++i) {
        auto Hc = [&](int a, int b) -> std::complex<double>& { return H[i][(size_t)a * m + b]; };
        for (int q = 0; q <= j; ++q)
          Hc(q, j) = std::complex<double>(h[i][q].x + h2w[i][q].x, h[i][q].y + h2w[i][q].y);
        Hc(j + 1, j) = hn[i];
        for (int q = 0; q < j; ++q) {
          const auto c = cs[i][q], sn_ = sn[i][q];
          const auto a = Hc(q, j), bb = Hc(q + 1, j);
          Hc(q, j) = c * a + sn_ * bb;
          Hc(q + 1, j) = -std::conj(sn_) * a + c * bb;
        }
        const auto h1 = Hc(j, j), h2v = Hc(j + 1, j);
        const double t = std::sqrt(std::norm(h1) + std::norm(h2v));
        std::complex<double> cj = 0, sj = 1;
        if (t > 0 && std::abs(h1) > 0) {
          cj = std::abs(h1) / t;
          sj = (h1 / std::abs(h1)) * std::conj(h2v) / t;
        }
        cs[i][j] = cj;
        sn[i][j] = sj;
        Hc(j, j) = cj * h1 + sj * h2v;
        Hc(j + 1, j) = 0;
        g[i][j + 1] = -std::conj(sj) * g[i][j];
        g[i][j] = cj * g[i][j];
        const double est = std::abs(g[i][j + 1]);
        worst_est = std::max(worst_est, bn[i] > 0 ? est / bn[i] : est);
      }
      k = j + 1;
      if (worst_est <= 0.5 * tol) break;
    }
    // x_i += Z_i y_i
    std::vector<std::vector<double2>> y(n);
    std::vector<int> nvk(n, k);
    for (int i = 0; i < n; ++i) {
      std::vector<std::complex<double>> yy(k, 0.0);
      for (int a = k - 1; a >= 0; --a) {
        std::complex<double> acc = g[i][a];
        for (int b = a + 1; b < k; ++b) acc -= H[i][(size_t)a * m + b] * yy[b];
        const auto d = H[i][(size_t)a * m + a];
        yy[a] = std::abs(d) > 0 ? acc / d : 0.0;
      }
      y[i].resize(k);
      for (int a = 0; a < k; ++a) y[i][a] = make_double2(yy[a].real(), yy[a].imag());
    }
    // lincomb works on V; point it at Z for this update
    std::vector<double2> one(n, make_double2(1, 0));
    std::vector<double2> hc((size_t)n * (m + 1), make_double2(0, 0));
    for (int i = 0; i < n; ++i)
      for (int q = 0; q < k; ++q) hc[(size_t)i * (m + 1) + q] = y[i][q];
    h2d(dcoef, hc.data(), hc.size() * 16, s);
    h2d(dbeta, one.data(), n * 16, s);
    {
      std::vector<double2*> xs(n);
      for (int i = 0; i < n; ++i) xs[i] = v[i].x;
      lincomb_dev(xs, nvk, true);
    }
  }
  HIP_OK(hipStreamSynchronize(s));
  return R;
}

}  // namespace aifmm
