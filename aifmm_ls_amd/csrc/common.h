// Shared device/host definitions for the MI355X (gfx950) LS solver library.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <stdexcept>
#include <string>

namespace aifmm {

#define HIP_OK(expr)                                                                          \
  do {                                                                                        \
    hipError_t _e = (expr);                                                                   \
    if (_e != hipSuccess)                                                                     \
      throw std::runtime_error(std::string("HIP error ") + hipGetErrorString(_e) + " at " +  \
                               __FILE__ + ":" + std::to_string(__LINE__) + ": " #expr);      \
  } while (0)

// ------------------------------------------------------------ complex ops
template <class T>
struct CplxOf;
template <>
struct CplxOf<double2> {
  using real = double;
};
template <>
struct CplxOf<float2> {
  using real = float;
};

__host__ __device__ inline double2 mk(double a, double b) { return make_double2(a, b); }
__host__ __device__ inline double2 cadd(double2 a, double2 b) { return mk(a.x + b.x, a.y + b.y); }
__host__ __device__ inline double2 csub(double2 a, double2 b) { return mk(a.x - b.x, a.y - b.y); }
__host__ __device__ inline double2 cmul(double2 a, double2 b) {
  return mk(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__host__ __device__ inline double2 cconj(double2 a) { return mk(a.x, -a.y); }
__host__ __device__ inline double2 cscale(double2 a, double s) { return mk(a.x * s, a.y * s); }
__host__ __device__ inline double cabs2(double2 a) { return a.x * a.x + a.y * a.y; }

__host__ __device__ inline float2 mkf(float a, float b) { return make_float2(a, b); }
__host__ __device__ inline float2 cadd(float2 a, float2 b) { return mkf(a.x + b.x, a.y + b.y); }
__host__ __device__ inline float2 csub(float2 a, float2 b) { return mkf(a.x - b.x, a.y - b.y); }
__host__ __device__ inline float2 cmul(float2 a, float2 b) {
  return mkf(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__host__ __device__ inline float2 cconj(float2 a) { return mkf(a.x, -a.y); }
__host__ __device__ inline float cabs2(float2 a) { return a.x * a.x + a.y * a.y; }

// Quadtree box index <-> integer coordinates.  Child c of a box sits at
// 0=(-,-) 1=(+,-) 2=(+,+) 3=(-,+) (LS_Assembly.hpp:834-849).
__host__ __device__ inline void box_xy(int level, int b, int& ix, int& iy) {
  ix = iy = 0;
  for (int t = level - 1; t >= 0; --t) {
    int c = (b >> (2 * t)) & 3;
    ix = 2 * ix + ((c == 1 || c == 2) ? 1 : 0);
    iy = 2 * iy + ((c == 2 || c == 3) ? 1 : 0);
  }
}
__host__ __device__ inline int xy_box(int level, int ix, int iy) {
  int b = 0;
  for (int t = level - 1; t >= 0; --t) {
    int bx = (ix >> t) & 1, by = (iy >> t) & 1;
    b = 4 * b + (by ? (bx ? 2 : 3) : (bx ? 1 : 0));
  }
  return b;
}

constexpr int SERIES = 60;  // Graf series length (FarFieldInteraction.hpp:55)
constexpr int NTERM = 2 * SERIES + 1;

// Everything a kernel needs to evaluate A(i,j) on the device.
struct EntryCtx {
  const double* px;
  const double* py;
  const double* q;
  const double2* M;  // NTERM x R, column-major (row = n + SERIES)
  const double2* T;  // 9 x R x R, column-major per neighbour slot
  int R;             // points per leaf
  int L;             // leaf level
  double kappa;
  double rleaf;
  double Ldom;
  // LS2DTree::findPhi (LS_Assembly.hpp:507-512, 530-535): 1, entries of the LS
  // operator I + kappa^2 q G; 0, of the volume potential G alone
  int find_phi = 1;
  // Adaptive discretisation (adaptive = 1; the fields above then leave M, T,
  // L and rleaf unused): leaf k at level leaf_lv[k], centre (lcx, lcy)[k];
  // rad[j] the box radius at level j; Mlev the far table of level j at
  // j * NTERM * R; the near leaves of leaf k are near_leaf[near_off[k] ..
  // near_off[k + 1]) with their tables Ttab + near_tab[..] * R * R.
  int adaptive = 0;
  const int* leaf_lv = nullptr;
  const double* lcx = nullptr;
  const double* lcy = nullptr;
  const double* rad = nullptr;
  const double2* Mlev = nullptr;
  const int* near_off = nullptr;
  const int* near_leaf = nullptr;
  const int* near_tab = nullptr;
  const double2* Ttab = nullptr;
};

}  // namespace aifmm
