// Common body of the source-level FMM2DTree stand-ins: the phase methods of
// the reference's FMM2DTree<kerneltype> (FMM2DTree_find_error.hpp:76-125 and
// the phases the drivers call, testFMM2D_find_error.cpp:126-166,
// testFMM2D_gen_rhs_x.cpp:96-133).  The phases map onto the C ABI:
//   createTree, assign_Tree_Interactions, assignLeafChargeLocations
//       -> checks only (the uniform tree and point order are fixed at
//          userkernel construction and equal the reference's)
//   assign_Leaf_rhs                      -> stages `properties` as the RHS
//   getNodes, assemble_M2L, initialise_phase, initialise_P2P_Leaf_Level
//       -> aifmm_ls_assemble (first call; the others are no-ops)
//   eliminate_phase_efficient / eliminate_phase -> aifmm_ls_factor
//   back_substitution_phase              -> aifmm_ls_solve
// error_check differs between the drivers: FMM2DTree.hpp (find_error) and
// FMM2DTree_gen_rhs_x.hpp (generator) each define FMM2DTree on this base, as
// the reference's two headers each define it; a program includes one.
// The factorisation does not depend on the RHS (the reference fuses its
// RHS sweep into the elimination; here it runs in back_substitution_phase).
// Extensions: `precision` (0 complex128 as the reference, 1 complex64) and
// `max_refine` / `refine_tol` (0: the reference's single direct solve).
#pragma once
#include <chrono>
#include <cmath>
#include <fstream>
#include <string>
#include <vector>

#include "cereal_bin.hpp"
#include "kernel.hpp"

template <class kerneltype>
class FMM2DTreeBase {
 public:
  kerneltype* K;
  int N, nLevels, TOL_POW;
  double L = 1.0;
  double* locations;
  dtype* properties;
  std::vector<int> boxNumbers, NumberOfParticlesInLeaves;
  // extensions
  int precision = 0;
  int max_refine = 0;
  double refine_tol = 1e-12;
  double solve_stats[8] = {0};
  std::vector<dtype> x;  // solution after back_substitution_phase

  FMM2DTreeBase(kerneltype* K, int N, int nLevels, int TOL_POW, double* locations, dtype* properties,
                std::vector<int>& boxNumbers, std::vector<int>& NumberOfParticlesInLeaves)
      : K(K), N(N), nLevels(nLevels), TOL_POW(TOL_POW), locations(locations), properties(properties),
        boxNumbers(boxNumbers), NumberOfParticlesInLeaves(NumberOfParticlesInLeaves) {
    if (N != K->N) throw std::runtime_error("FMM2DTree: N differs from the kernel's");
    aifmm::check(aifmm_ls_set_tolerance(K->handle(), TOL_POW));
  }

  void createTree() {
    if ((int)boxNumbers.size() != (1 << (2 * nLevels)))
      throw std::runtime_error("FMM2DTree: boxNumbers must list the 4^nLevels leaves of the uniform tree");
  }
  void assign_Tree_Interactions() {}
  void assignLeafChargeLocations() {
    for (int i = 0; i < N; ++i)
      if (locations[2 * (size_t)i] != K->gridPoints[i].x || locations[2 * (size_t)i + 1] != K->gridPoints[i].y)
        throw std::runtime_error("FMM2DTree: locations must be the kernel's gridPoints in order");
  }
  void assign_Leaf_rhs() {
    b_.assign(2 * (size_t)N, 0.0);
    for (int i = 0; i < N; ++i) b_[2 * (size_t)i] = properties[i].real(), b_[2 * (size_t)i + 1] = properties[i].imag();
  }
  void getNodes() { assemble(); }
  void assemble_M2L() { assemble(); }
  void initialise_phase() { assemble(); }
  void initialise_P2P_Leaf_Level() { assemble(); }
  void eliminate_phase_efficient() { aifmm::check(aifmm_ls_factor(K->handle(), precision)); }
  void eliminate_phase() { eliminate_phase_efficient(); }
  void back_substitution_phase() {
    if (b_.empty()) assign_Leaf_rhs();
    std::vector<double> xx(2 * (size_t)N);
    aifmm::check(aifmm_ls_solve(K->handle(), 1, b_.data(), xx.data(), refine_tol, max_refine, solve_stats));
    x.resize(N);
    for (int i = 0; i < N; ++i) x[i] = dtype(xx[2 * (size_t)i], xx[2 * (size_t)i + 1]);
  }

 protected:
  std::vector<double> b_;  // the staged RHS, interleaved
  bool assembled_ = false;
  void assemble() {
    if (assembled_) return;
    aifmm::check(aifmm_ls_assemble(K->handle()));
    assembled_ = true;
  }
  // rhs_x/<what>_{real,imag}_{Qchoice}_{int kappa}_{N}.bin (the reference's names)
  std::string rhs_x_name(const std::string& what, const std::string& part) const {
    return "rhs_x/" + what + "_" + part + "_" + std::to_string(Qchoice) + "_" + std::to_string((int)kappa) + "_" +
           std::to_string(N) + ".bin";
  }
  std::vector<double> interleaved_x() const {
    std::vector<double> xx(2 * (size_t)N);
    for (int i = 0; i < N; ++i) xx[2 * (size_t)i] = x[i].real(), xx[2 * (size_t)i + 1] = x[i].imag();
    return xx;
  }
};
