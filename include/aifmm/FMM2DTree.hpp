// Source-level stand-in for the reference's FMM2DTree<kerneltype> of the
// find_error driver (FMM2DTree_find_error.hpp:76-2583): the phase methods of
// FMM2DTree_base.hpp, and error_check as the reference's (:2534-2581):
// forward error against rhs_x/x_{real,imag}_{Qchoice}_{int kappa}_{N}.bin when
// present (as the reference), else the relative residual |b - A_H x| / |b|.
#pragma once
#include "FMM2DTree_base.hpp"

template <class kerneltype>
class FMM2DTree : public FMM2DTreeBase<kerneltype> {
  using Base = FMM2DTreeBase<kerneltype>;

 public:
  using Base::Base;
  using Base::N;
  using Base::x;

  // FMM2DTree_find_error.hpp:2534-2581
  double error_check() {
    const std::string fr = this->rhs_x_name("x", "real"), fi = this->rhs_x_name("x", "imag");
    if (std::ifstream(fr).good() && std::ifstream(fi).good()) {
      std::vector<double> re = aifmm::load_vec(fr), im = aifmm::load_vec(fi);
      if ((int)re.size() != N || (int)im.size() != N) throw std::runtime_error("x_true has the wrong length");
      double e = 0, n = 0;
      for (int i = 0; i < N; ++i) {
        dtype t(re[i], im[i]);
        e += std::norm(x[i] - t);
        n += std::norm(t);
      }
      return std::sqrt(e / n);
    }
    std::vector<double> xx = this->interleaved_x(), ax(2 * (size_t)N);
    aifmm::check(aifmm_ls_matvec(this->K->handle(), 1, xx.data(), ax.data()));
    const std::vector<double>& b = this->b_;
    double e = 0, n = 0;
    for (size_t t = 0; t < b.size(); ++t) e += (b[t] - ax[t]) * (b[t] - ax[t]), n += b[t] * b[t];
    return std::sqrt(e / n);
  }
};
