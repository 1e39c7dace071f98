// Prints what the source-compatible userkernel (include/aifmm/kernel.hpp)
// returns for getMatrixEntry, getRow, getCol and getMatrix on the reference's
// kernel at 2 levels (kappa 10, Qchoice 0), for tests/test_shim.py to check
// against the oracle's entries.  One "i j re im" line per value.
#include <cstdio>
#include <vector>

#include "aifmm/kernel.hpp"

static void out(int i, int j, dtype v) { std::printf("%d %d %.17g %.17g\n", i, j, v.real(), v.imag()); }

int main() {
  kappa = 10.0;
  Qchoice = 0;
  std::vector<pts2D> px, py;
  userkernel K(6, 1, 6, 1.0, px, py, 2);
  std::printf("N %d %d\n", K.N, K.nLevelsResolve);
  const int pairs[][2] = {{0, 0}, {5, 100}, {100, 5}, {300, 575}, {575, 0}, {36, 37}, {200, 200}};
  for (auto& p : pairs) out(p[0], p[1], K.getMatrixEntry(p[0], p[1]));
  const std::vector<int> cols = {0, 1, 2, 50, 200, 575};
  Vec r = K.getRow(7, cols);
  for (size_t t = 0; t < cols.size(); ++t) out(7, cols[t], r(t));
  const std::vector<int> rows = {3, 4, 400};
  Vec c = K.getCol(9, rows);
  for (size_t t = 0; t < rows.size(); ++t) out(rows[t], 9, c(t));
  const std::vector<int> mr = {1, 2, 3}, mc = {40, 41};
  Mat M = K.getMatrix(mr, mc);
  for (size_t b = 0; b < mc.size(); ++b)
    for (size_t a = 0; a < mr.size(); ++a) out(mr[a], mc[b], M((int)a, (int)b));
  return 0;
}
