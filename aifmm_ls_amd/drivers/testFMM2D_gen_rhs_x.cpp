// The generator driver (testFMM2D_gen_rhs_x.cpp:14-138) on the MI355X solver,
// through the source-compatible FMM2DTree (generator variant) / userkernel
// headers.
//
//   testFMM2D_gen_rhs_x nChebNodes kappa degree L Qchoice TOL_POW nLevels
//                       [precision max_refine refine_tol]
//
// Solves for the kernel's incident-field RHS, then writes the dense
// rhs = A x and x to rhs_x/{rhs,x}_{real,imag}_{Qchoice}_{int kappa}_{N}.bin
// (the files the find_error driver reads) and prints |A x - b| / |b|.
// precision 0 = complex128 factors (the reference), 1 = complex64; max_refine
// > 0 refines against the FP64 operator to refine_tol.
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <iostream>
#include <string>

#include "aifmm/FMM2DTree_gen_rhs_x.hpp"

static double now() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

int main(int argc, char* argv[]) {
  if (argc < 8) {
    std::fprintf(stderr, "usage: %s nChebNodes kappa degree L Qchoice TOL_POW nLevels [precision max_refine refine_tol]\n",
                 argv[0]);
    return 2;
  }
  try {
    const int nChebNodes = atoi(argv[1]);
    kappa = atof(argv[2]);
    const int degreeOfBases = atoi(argv[3]);
    const double L = atof(argv[4]);
    Qchoice = atoi(argv[5]);
    const int TOL_POW = atoi(argv[6]);
    const int nLevelsUniform = atoi(argv[7]);
    const int precision = argc > 8 ? atoi(argv[8]) : 0;
    const int max_refine = argc > 9 ? atoi(argv[9]) : 0;
    const double refine_tol = argc > 10 ? atof(argv[10]) : 1e-12;
    std::cout << "kappa: " << kappa << "\nnChebNodes: " << nChebNodes << "\nL: " << L << "\nQchoice: " << Qchoice
              << "\nTOL_POW: " << TOL_POW << "\nnLevelsUniform: " << nLevelsUniform << std::endl;

    std::vector<pts2D> px, py;
    userkernel* K = new userkernel(nChebNodes, 1, degreeOfBases, L, px, py, nLevelsUniform);
    const int N = K->N;
    const int nLevels = K->nLevelsResolve;
    double* locations = new double[2 * (size_t)N];
    dtype* properties = new dtype[N];
    for (int j = 0; j < N; ++j) {
      locations[2 * (size_t)j] = K->gridPoints[j].x;
      locations[2 * (size_t)j + 1] = K->gridPoints[j].y;
      properties[j] = K->rhs(j);
    }
    std::vector<int> boxNumbers, NumberOfParticlesInLeaves;
    for (int i = 0; i < (1 << (2 * nLevels)); ++i) {
      boxNumbers.push_back(i);
      NumberOfParticlesInLeaves.push_back(nChebNodes * nChebNodes);
    }
    std::cout << "N: " << N << std::endl;
    FMM2DTree<userkernel>* A = new FMM2DTree<userkernel>(K, N, nLevels, TOL_POW, locations, properties, boxNumbers,
                                                         NumberOfParticlesInLeaves);
    A->precision = precision;
    A->max_refine = max_refine;
    A->refine_tol = refine_tol;

    double s = now();
    A->createTree();
    A->assign_Tree_Interactions();
    std::cout << "Time taken to create the tree is: " << now() - s << std::endl;
    s = now();
    A->assignLeafChargeLocations();
    A->assign_Leaf_rhs();
    A->getNodes();
    A->assemble_M2L();
    A->initialise_phase();
    A->initialise_P2P_Leaf_Level();
    std::cout << "Time taken to assemble is: " << now() - s << std::endl;
    s = now();
    A->eliminate_phase_efficient();
    std::cout << "Time taken to factorize (eliminate phase): " << now() - s << std::endl;
    s = now();
    A->back_substitution_phase();
    std::cout << "Time taken to solve (back substitution phase): " << now() - s << std::endl;
    std::cout << "Performing error calculation... " << std::endl;
    A->assign_Leaf_rhs();
    s = now();
    const double error = A->error_check();
    std::cout << "Time taken by the dense product: " << now() - s << std::endl;
    std::cout << "error: " << error << std::endl;
    delete A;
    delete[] locations;
    delete[] properties;
    delete K;
    return 0;
  } catch (const std::exception& e) {
    std::fprintf(stderr, "error: %s\n", e.what());
    return 1;
  }
}
