// Source-level stand-in for the reference's kernel.hpp / LS_Assembly.hpp
// surface (kernel.hpp:25-31, 90-193; LS_Assembly.hpp:31-38): the globals
// kappa and Qchoice, pts2D, dtype, and userkernel with getMatrixEntry,
// getRow, getCol and getMatrix.  Every call goes through the C ABI of
// libaifmm_ls.so (include/aifmm_ls.h); the entries are evaluated on the GPU.
//
// Differences a caller sees: Vec and Mat are small std-library containers
// (no Eigen in this build); Qchoice 6 selects a sum of Gaussian bumps set
// with aifmm::set_contrast_bumps; errors throw std::runtime_error instead of
// aborting.
#pragma once
#include <cmath>
#include <complex>
#include <cstdint>
#include <cstdlib>
#include <stdexcept>
#include <string>
#include <vector>

#include "../aifmm_ls.h"

using dtype = std::complex<double>;

struct pts2D {
  double x, y;
};

// the reference's globals (LS_Assembly.hpp:31-35), read when a kernel is built
inline double kappa = 0.0;
inline int Qchoice = 0;

// Column vector and column-major matrix of dtype.
struct Vec {
  std::vector<dtype> a;
  Vec() {}
  explicit Vec(size_t n) : a(n) {}
  size_t size() const { return a.size(); }
  dtype& operator()(size_t i) { return a[i]; }
  const dtype& operator()(size_t i) const { return a[i]; }
  double norm() const {
    double s = 0;
    for (const auto& v : a) s += std::norm(v);
    return std::sqrt(s);
  }
};

struct Mat {
  int r = 0, c = 0;
  std::vector<dtype> a;
  Mat() {}
  Mat(int r_, int c_) : r(r_), c(c_), a((size_t)r_ * c_) {}
  int rows() const { return r; }
  int cols() const { return c; }
  dtype& operator()(int i, int j) { return a[i + (size_t)r * j]; }
  const dtype& operator()(int i, int j) const { return a[i + (size_t)r * j]; }
};

namespace aifmm {

inline void check(int rc) {
  if (rc != 0) throw std::runtime_error(aifmm_ls_last_error());
}

// Contrast for Qchoice 6: q(x) = sum_m a_m exp(-|x - c_m|^2 / w_m^2); rows
// (a, w, cx, cy).
inline std::vector<double>& contrast_bumps() {
  static std::vector<double> b;
  return b;
}
inline void set_contrast_bumps(const std::vector<double>& rows_a_w_cx_cy) { contrast_bumps() = rows_a_w_cx_cy; }

// Where the tables come from (LS_Assembly.hpp:262-275).  PRECOMP_COMPUTE: GPU
// quadrature, no files.  PRECOMP_WRITE: the same, and every table written as
// precomputations/ under precomp_dir(), as the reference's ctor does.
// PRECOMP_READ: read from precomputations/ under precomp_dir(), as the
// reference does with its three write lines commented out (LS/Readme.md:1-5).
// The default comes from the environment variable AIFMM_PRECOMPUTATIONS
// ("write" or "read"; anything else: compute) and the directory from
// AIFMM_PRECOMPUTATIONS_DIR (default: the working directory).
enum { PRECOMP_COMPUTE = 0, PRECOMP_WRITE = 1, PRECOMP_READ = 2 };
inline int& precomp_mode() {
  static int m = [] {
    const char* e = std::getenv("AIFMM_PRECOMPUTATIONS");
    if (!e) return (int)PRECOMP_COMPUTE;
    std::string v(e);
    return v == "write" ? (int)PRECOMP_WRITE : v == "read" ? (int)PRECOMP_READ : (int)PRECOMP_COMPUTE;
  }();
  return m;
}
inline std::string& precomp_dir() {
  static std::string d = [] {
    const char* e = std::getenv("AIFMM_PRECOMPUTATIONS_DIR");
    return std::string(e && *e ? e : ".");
  }();
  return d;
}

}  // namespace aifmm

class userkernel {
 public:
  int N = 0;
  int nLevelsResolve = 0;
  std::vector<pts2D> gridPoints;
  Vec rhs;  // incident-field RHS, LS_functions.hpp:113-120 (theta = 0)
  std::vector<pts2D> particles_X, particles_Y;
  // LS2DTree::findPhi (LS_Assembly.hpp:195): true, getMatrixEntry and friends
  // return entries of I + kappa^2 q G; false, of the volume potential G
  bool findPhi = true;

  // kernel.hpp:106-128.  treeAdaptivity and the particle lists are accepted
  // for source compatibility (uniform tree only).
  userkernel(int nChebNodes, int treeAdaptivity, int degreeOfBases, double L, std::vector<pts2D>& particles_X,
             std::vector<pts2D>& particles_Y, int nLevelsUniform)
      : particles_X(particles_X), particles_Y(particles_Y) {
    (void)treeAdaptivity;
    aifmm_ls_params p{};
    p.n_cheb = nChebNodes;
    p.degree = degreeOfBases;
    p.n_levels = nLevelsUniform;
    p.tol_pow = 9;  // the kernel's own value (kernel.hpp:117); FMM2DTree sets the one in use
    p.qchoice = Qchoice;
    p.quantize_tables = 1;
    p.kappa = kappa;
    p.L = L;
    const auto& bumps = aifmm::contrast_bumps();
    p.bumps = bumps.empty() ? nullptr : bumps.data();
    p.n_bumps = (int)(bumps.size() / 4);
    if (aifmm::precomp_mode() == aifmm::PRECOMP_READ) {
      aifmm::check(aifmm_ls_create_cached(&p, aifmm::precomp_dir().c_str(), &h_));
    } else {
      aifmm::check(aifmm_ls_create(&p, &h_));
      if (aifmm::precomp_mode() == aifmm::PRECOMP_WRITE)
        aifmm::check(aifmm_ls_write_precomputations(h_, aifmm::precomp_dir().c_str(), nullptr));
    }
    N = (int)aifmm_ls_size(h_);
    nLevelsResolve = nLevelsUniform;
    std::vector<double> xy(2 * (size_t)N), q(N);
    aifmm::check(aifmm_ls_points(h_, xy.data(), q.data()));
    gridPoints.resize(N);
    for (int i = 0; i < N; ++i) gridPoints[i] = pts2D{xy[2 * (size_t)i], xy[2 * (size_t)i + 1]};
    std::vector<double> b(2 * (size_t)N);
    aifmm::check(aifmm_ls_rhs_planewave(h_, 0.0, b.data()));
    rhs = Vec(N);
    for (int i = 0; i < N; ++i) rhs(i) = dtype(b[2 * (size_t)i], b[2 * (size_t)i + 1]);
  }
  ~userkernel() {
    if (h_) aifmm_ls_destroy(h_);
  }
  userkernel(const userkernel&) = delete;
  userkernel& operator=(const userkernel&) = delete;

  // kernel.hpp:130-136
  dtype getMatrixEntry(const unsigned i, const unsigned j) {
    int64_t I = i, J = j;
    double v[2];
    aifmm::check(aifmm_ls_set_find_phi(h_, findPhi ? 1 : 0));
    aifmm::check(aifmm_ls_entries(h_, 1, &I, &J, v));
    return dtype(v[0], v[1]);
  }
  // kernel.hpp:54-86: row j over the columns in col_indices, and so on
  Vec getRow(const int j, const std::vector<int>& col_indices) {
    std::vector<int64_t> I(col_indices.size(), j), J(col_indices.begin(), col_indices.end());
    return gather(I, J);
  }
  Vec getCol(const int k, const std::vector<int>& row_indices) {
    std::vector<int64_t> I(row_indices.begin(), row_indices.end()), J(row_indices.size(), k);
    return gather(I, J);
  }
  Mat getMatrix(const std::vector<int>& row_indices, const std::vector<int>& col_indices) {
    Mat M((int)row_indices.size(), (int)col_indices.size());
    std::vector<int64_t> I, J;
    I.reserve(M.a.size());
    J.reserve(M.a.size());
    for (int c : col_indices)
      for (int r : row_indices) I.push_back(r), J.push_back(c);
    Vec v = gather(I, J);
    M.a = std::move(v.a);
    return M;
  }

  // the underlying C ABI handle
  aifmm_ls_problem* handle() const { return h_; }

 private:
  aifmm_ls_problem* h_ = nullptr;
  Vec gather(const std::vector<int64_t>& I, const std::vector<int64_t>& J) {
    Vec out(I.size());
    if (I.empty()) return out;
    std::vector<double> v(2 * I.size());
    aifmm::check(aifmm_ls_set_find_phi(h_, findPhi ? 1 : 0));
    aifmm::check(aifmm_ls_entries(h_, (int64_t)I.size(), I.data(), J.data(), v.data()));
    for (size_t t = 0; t < I.size(); ++t) out(t) = dtype(v[2 * t], v[2 * t + 1]);
    return out;
  }
};
