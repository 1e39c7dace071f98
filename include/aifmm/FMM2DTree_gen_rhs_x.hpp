// Source-level stand-in for the reference's FMM2DTree<kerneltype> of the
// generator driver (FMM2DTree_gen_rhs_x.hpp): the phase methods of
// FMM2DTree_base.hpp, and error_check as the generator's (:2808-2903): the
// dense rhs = A x of the solution x over all leaf pairs (aifmm_ls_dense_matvec,
// every entry as getMatrixEntry), rhs and x written to
// rhs_x/{rhs,x}_{real,imag}_{Qchoice}_{int kappa}_{N}.bin, and the relative
// residual |A x - b| / |b| against the staged right-hand side b returned.
// The reference prints the max-norms of A x - b and b; so does this.
#pragma once
#include <filesystem>
#include <iostream>

#include "FMM2DTree_base.hpp"

template <class kerneltype>
class FMM2DTree : public FMM2DTreeBase<kerneltype> {
  using Base = FMM2DTreeBase<kerneltype>;

 public:
  using Base::Base;
  using Base::N;
  using Base::x;

  // FMM2DTree_gen_rhs_x.hpp:2808-2903
  double error_check() {
    std::vector<double> xx = this->interleaved_x(), ax(2 * (size_t)N);
    aifmm::check(aifmm_ls_dense_matvec(this->K->handle(), 1, xx.data(), ax.data()));
    std::filesystem::create_directories("rhs_x");
    std::vector<double> re(N), im(N);
    for (int i = 0; i < N; ++i) re[i] = ax[2 * (size_t)i], im[i] = ax[2 * (size_t)i + 1];
    aifmm::save_vec(this->rhs_x_name("rhs", "real"), re);
    aifmm::save_vec(this->rhs_x_name("rhs", "imag"), im);
    for (int i = 0; i < N; ++i) re[i] = x[i].real(), im[i] = x[i].imag();
    aifmm::save_vec(this->rhs_x_name("x", "real"), re);
    aifmm::save_vec(this->rhs_x_name("x", "imag"), im);
    const std::vector<double>& b = this->b_;
    double e2 = 0, b2 = 0, einf = 0, binf = 0;
    for (int i = 0; i < N; ++i) {
      const double er = ax[2 * (size_t)i] - b[2 * (size_t)i], ei = ax[2 * (size_t)i + 1] - b[2 * (size_t)i + 1];
      const double br = b[2 * (size_t)i], bi = b[2 * (size_t)i + 1];
      e2 += er * er + ei * ei;
      b2 += br * br + bi * bi;
      einf = std::max(einf, std::hypot(er, ei));
      binf = std::max(binf, std::hypot(br, bi));
    }
    std::cout << "gen rhs x" << std::endl;
    std::cout << "abs error: " << einf << std::endl;
    std::cout << "b norm: " << binf << std::endl;
    return std::sqrt(e2 / b2);
  }
};
