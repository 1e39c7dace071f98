// ORACLE — test infrastructure only (see ls_oracle.hpp).
// Dense complex linear algebra for the CPU restatement: products, LU with
// partial pivoting, and column-pivoted Householder QR with the semantics of
// Eigen::ColPivHouseholderQR that the reference relies on.
#include <algorithm>
#include <cfloat>
#include <cmath>

#include "ls_oracle.hpp"

namespace orc {

Mat mul(const Mat& A, const Mat& B) {
  Mat C(A.r, B.c);
  for (int j = 0; j < B.c; ++j)
    for (int k = 0; k < A.c; ++k) {
      cd b = B(k, j);
      if (b == cd(0, 0)) continue;
      const cd* a = &A.a[(size_t)k * A.r];
      cd* c = &C.a[(size_t)j * C.r];
      for (int i = 0; i < A.r; ++i) c[i] += a[i] * b;
    }
  return C;
}

Mat mulAH(const Mat& A, const Mat& B) {
  Mat C(A.c, B.c);
  for (int j = 0; j < B.c; ++j)
    for (int i = 0; i < A.c; ++i) {
      cd s = 0;
      const cd* a = &A.a[(size_t)i * A.r];
      const cd* b = &B.a[(size_t)j * B.r];
      for (int k = 0; k < A.r; ++k) s += std::conj(a[k]) * b[k];
      C(i, j) = s;
    }
  return C;
}

Mat adj(const Mat& A) {
  Mat C(A.c, A.r);
  for (int j = 0; j < A.c; ++j)
    for (int i = 0; i < A.r; ++i) C(j, i) = std::conj(A(i, j));
  return C;
}

Mat neg(const Mat& A) {
  Mat C = A;
  for (auto& v : C.a) v = -v;
  return C;
}

Mat block(const Mat& A, int i0, int j0, int nr, int nc) {
  Mat C(nr, nc);
  for (int j = 0; j < nc; ++j)
    for (int i = 0; i < nr; ++i) C(i, j) = A(i0 + i, j0 + j);
  return C;
}

Mat hcat(const Mat& A, const Mat& B) {
  // An empty block contributes nothing (Eigen comma initializer behaviour).
  if (B.size() == 0) return A;
  if (A.size() == 0 && A.c == 0) {
    Mat C(B.r, B.c);
    C.a = B.a;
    return C;
  }
  Mat C(A.r, A.c + B.c);
  std::copy(A.a.begin(), A.a.end(), C.a.begin());
  std::copy(B.a.begin(), B.a.end(), C.a.begin() + A.size());
  return C;
}

Mat eye(int n) {
  Mat I(n, n);
  for (int i = 0; i < n; ++i) I(i, i) = 1.0;
  return I;
}

double fro(const Mat& A) {
  double s = 0;
  for (auto& v : A.a) s += std::norm(v);
  return std::sqrt(s);
}

void subeq(Mat& A, const Mat& B) {
  if (A.size() == 0 && B.size() != 0) {
    A = neg(B);
    return;
  }
  if (B.size() == 0) return;
  for (size_t i = 0; i < A.a.size(); ++i) A.a[i] -= B.a[i];
}

// ------------------------------------------------------------------ LU
void LU::compute(const Mat& A) {
  n = A.r;
  f = A;
  piv.assign(n, 0);
  for (int k = 0; k < n; ++k) {
    int p = k;
    double best = std::abs(f(k, k));
    for (int i = k + 1; i < n; ++i)
      if (std::abs(f(i, k)) > best) best = std::abs(f(i, k)), p = i;
    piv[k] = p;
    if (p != k)
      for (int j = 0; j < n; ++j) std::swap(f(k, j), f(p, j));
    cd d = f(k, k);
    if (d == cd(0, 0)) continue;
    for (int i = k + 1; i < n; ++i) f(i, k) /= d;
    for (int j = k + 1; j < n; ++j) {
      cd u = f(k, j);
      if (u == cd(0, 0)) continue;
      for (int i = k + 1; i < n; ++i) f(i, j) -= f(i, k) * u;
    }
  }
}

Mat LU::solve(const Mat& B) const {
  Mat X = B;
  for (int j = 0; j < X.c; ++j) {
    for (int k = 0; k < n; ++k)
      if (piv[k] != k) std::swap(X(k, j), X(piv[k], j));
    for (int k = 0; k < n; ++k)
      for (int i = k + 1; i < n; ++i) X(i, j) -= f(i, k) * X(k, j);
    for (int k = n - 1; k >= 0; --k) {
      X(k, j) /= f(k, k);
      for (int i = 0; i < k; ++i) X(i, j) -= f(i, k) * X(k, j);
    }
  }
  return X;
}

// ------------------------------------------------------------------ CPQR
void CPQR::compute(const Mat& A, double nz_eps) {
  m = A.r;
  n = A.c;
  qr = A;
  int size = std::min(m, n);
  tau.assign(size, cd(0, 0));
  perm.resize(n);
  for (int j = 0; j < n; ++j) perm[j] = j;
  std::vector<double> upd(n), dir(n);
  double maxnorm = 0;
  for (int j = 0; j < n; ++j) {
    double s = 0;
    for (int i = 0; i < m; ++i) s += std::norm(qr(i, j));
    upd[j] = dir[j] = std::sqrt(s);
    maxnorm = std::max(maxnorm, upd[j]);
  }
  double eps = nz_eps > 0 ? nz_eps : DBL_EPSILON;
  double th_helper = (maxnorm * eps) * (maxnorm * eps) / (m > 0 ? m : 1);
  double norm_downdate_threshold = std::sqrt(eps);
  nonzero = size;
  maxpivot = 0;
  for (int k = 0; k < size; ++k) {
    int big = k;
    for (int j = k + 1; j < n; ++j)
      if (upd[j] > upd[big]) big = j;
    double big_sq = upd[big] * upd[big];
    if (nonzero == size && big_sq < th_helper * (m - k)) nonzero = k;
    if (big != k) {
      for (int i = 0; i < m; ++i) std::swap(qr(i, k), qr(i, big));
      std::swap(upd[k], upd[big]);
      std::swap(dir[k], dir[big]);
      std::swap(perm[k], perm[big]);
    }
    // Householder reflector I - t v v^H mapping column k to beta e_1.
    cd c0 = qr(k, k);
    double tail = 0;
    for (int i = k + 1; i < m; ++i) tail += std::norm(qr(i, k));
    double beta;
    cd t;
    const double tiny = DBL_MIN;
    if (tail <= tiny && std::imag(c0) * std::imag(c0) <= tiny) {
      t = 0;
      beta = std::real(c0);
      for (int i = k + 1; i < m; ++i) qr(i, k) = 0;
    } else {
      beta = std::sqrt(std::norm(c0) + tail);
      if (std::real(c0) >= 0) beta = -beta;
      cd den = c0 - beta;
      for (int i = k + 1; i < m; ++i) qr(i, k) /= den;
      t = std::conj((beta - c0) / beta);
    }
    tau[k] = t;
    qr(k, k) = beta;
    maxpivot = std::max(maxpivot, std::abs(beta));
    // apply to the trailing columns: x -= t * v * (v^H x)
    for (int j = k + 1; j < n; ++j) {
      cd s = qr(k, j);
      for (int i = k + 1; i < m; ++i) s += std::conj(qr(i, k)) * qr(i, j);
      s *= t;
      qr(k, j) -= s;
      for (int i = k + 1; i < m; ++i) qr(i, j) -= s * qr(i, k);
    }
    for (int j = k + 1; j < n; ++j) {
      if (upd[j] != 0) {
        double temp = std::abs(qr(k, j)) / upd[j];
        temp = (1 + temp) * (1 - temp);
        if (temp < 0) temp = 0;
        double r = upd[j] / dir[j];
        double temp2 = temp * r * r;
        if (temp2 <= norm_downdate_threshold) {
          double s = 0;
          for (int i = k + 1; i < m; ++i) s += std::norm(qr(i, j));
          dir[j] = upd[j] = std::sqrt(s);
        } else {
          upd[j] *= std::sqrt(temp);
        }
      }
    }
  }
}

int CPQR::rank(double threshold) const {
  double pre = std::abs(maxpivot) * threshold;
  int r = 0;
  for (int i = 0; i < std::min(m, n); ++i)
    if (std::abs(qr(i, i)) > pre) ++r;
  return r;
}

Mat CPQR::solve(const Mat& B) const {
  Mat c = B;
  int nz = nonzero;
  for (int k = 0; k < nz; ++k) {
    cd t = tau[k];
    // apply (I - t v v^H)^H = I - conj(t) v v^H ... Q^H = H_{nz-1}..H_0 with
    // H_k = I - t_k v v^H being the reflector used in compute().
    for (int j = 0; j < c.c; ++j) {
      cd s = c(k, j);
      for (int i = k + 1; i < m; ++i) s += std::conj(qr(i, k)) * c(i, j);
      s *= t;
      c(k, j) -= s;
      for (int i = k + 1; i < m; ++i) c(i, j) -= s * qr(i, k);
    }
  }
  Mat X(n, B.c);
  for (int j = 0; j < B.c; ++j) {
    for (int i = nz - 1; i >= 0; --i) {
      cd s = c(i, j);
      for (int l = i + 1; l < nz; ++l) s -= qr(i, l) * c(l, j);
      c(i, j) = s / qr(i, i);
    }
    for (int i = 0; i < nz; ++i) X(perm[i], j) = c(i, j);
  }
  return X;
}

Mat CPQR::Q(int k) const {
  Mat E(m, k);
  for (int i = 0; i < std::min(m, k); ++i) E(i, i) = 1.0;
  int size = std::min(m, n);
  for (int h = size - 1; h >= 0; --h) {
    cd t = std::conj(tau[h]);  // Q = H_0^H H_1^H ...
    for (int j = 0; j < k; ++j) {
      cd s = E(h, j);
      for (int i = h + 1; i < m; ++i) s += std::conj(qr(i, h)) * E(i, j);
      s *= t;
      E(h, j) -= s;
      for (int i = h + 1; i < m; ++i) E(i, j) -= s * qr(i, h);
    }
  }
  return E;
}

}  // namespace orc
