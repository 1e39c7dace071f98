// ORACLE — test infrastructure only (see ls_oracle.hpp).
// Discretisation of the 2D Lippmann–Schwinger operator: Gauss–Legendre
// rule, Bessel functions, adaptive panel quadrature, the far-field (M) and
// colleague near-field (T) tables with the 6-significant-digit text round
// trip, contrast functions and the matrix entry A(i,j).
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <functional>
#include <stdexcept>

#include "ls_oracle.hpp"

namespace orc {

static const double PI = 3.1415926535897932384;  // LS_Assembly.hpp:30

// Gauss–Legendre nodes/weights by Newton iteration on P_n.  Ordered as the
// reference's table (LS/Gauss_Legendre_Nodes_and_Weights.hpp:210-245 for
// n=10): pairs (-x, +x) with |x| ascending, 0 first for odd n.
void gauss_legendre(int n, std::vector<double>& x, std::vector<double>& w) {
  std::vector<long double> pos, wpos;
  int half = n / 2;
  for (int i = 0; i < half; ++i) {
    // i-th positive root counted from the largest
    long double z = cosl(3.14159265358979323846264338327950288L * (i + 0.75L) / (n + 0.5L));
    long double dp = 0;
    for (int it = 0; it < 100; ++it) {
      long double p0 = 1, p1 = z;
      for (int k = 2; k <= n; ++k) {
        long double p2 = ((2 * k - 1) * z * p1 - (k - 1) * p0) / k;
        p0 = p1;
        p1 = p2;
      }
      dp = n * (z * p1 - p0) / (z * z - 1);
      long double dz = p1 / dp;
      z -= dz;
      if (fabsl(dz) < 1e-19L) break;
    }
    long double p0 = 1, p1 = z;
    for (int k = 2; k <= n; ++k) {
      long double p2 = ((2 * k - 1) * z * p1 - (k - 1) * p0) / k;
      p0 = p1;
      p1 = p2;
    }
    dp = n * (z * p1 - p0) / (z * z - 1);
    pos.push_back(z);
    wpos.push_back(2 / ((1 - z * z) * dp * dp));
  }
  x.clear();
  w.clear();
  if (n % 2 == 1) {
    long double p0 = 1, p1 = 0;
    for (int k = 2; k <= n; ++k) {
      long double p2 = (-(k - 1) * p0) / k;
      p0 = p1;
      p1 = p2;
    }
    // P_n'(0) = n P_{n-1}(0)
    long double dp = n * p0;
    x.push_back(0.0);
    w.push_back((double)(2 / (dp * dp)));
  }
  for (int i = half - 1; i >= 0; --i) {  // ascending |x|
    x.push_back((double)-pos[i]);
    w.push_back((double)wpos[i]);
    x.push_back((double)pos[i]);
    w.push_back((double)wpos[i]);
  }
}

void bessel_jy(double x, int nmax, double* J, double* Y, bool exact) {
  if (exact) {
    for (int n = 0; n <= nmax; ++n) {
      J[n] = std::cyl_bessel_j((double)n, x);
      Y[n] = std::cyl_neumann((double)n, x);
    }
    return;
  }
  Y[0] = ::y0(x);
  if (nmax >= 1) Y[1] = ::y1(x);
  for (int n = 1; n < nmax; ++n) Y[n + 1] = (2.0 * n / x) * Y[n] - Y[n - 1];
  // Miller backward recurrence from an even start index above max(nmax,x).
  int big = (int)std::max((double)nmax, x);
  int M = big + 30 + (int)std::sqrt(40.0 * big);
  if (M % 2) ++M;
  double jp1 = 0, jc = 1e-300, sum = 0;
  std::vector<double> tmp(nmax + 1, 0.0);
  double scale_acc = 1.0;  // not needed beyond rescaling inside loop
  (void)scale_acc;
  for (int k = M; k >= 1; --k) {
    double jm1 = (2.0 * k / x) * jc - jp1;
    jp1 = jc;
    jc = jm1;  // jc = j_{k-1}
    if (std::fabs(jc) > 1e250) {
      jc *= 1e-250;
      jp1 *= 1e-250;
      sum *= 1e-250;
      for (int n = 0; n <= nmax; ++n) tmp[n] *= 1e-250;
    }
    int km1 = k - 1;
    if (km1 <= nmax) tmp[km1] = jc;
    if (km1 > 0 && (km1 % 2) == 0) sum += 2 * jc;
  }
  sum += jc;  // j_0
  for (int n = 0; n <= nmax; ++n) J[n] = tmp[n] / sum;
}

// ---------------------------------------------------------- contrast (LS_functions.hpp)
static double q_gaussian(double x, double y) {  // LS_functions.hpp:1-4
  double R2 = x * x + y * y;
  return 1.5 * exp(-160.0 * R2);
}
static double q_multi(double x, double y) {  // LS_functions.hpp:6-52
  static const double cx[20] = {-1.17182014706864, -0.861066625441127, -1.14533939124057, -0.856639124568525,
                                -0.620066350133170, -0.457064031294315, -0.482782006538245, -0.635070123186999,
                                0.108594058866164, -0.0315722590552472, -0.177274727610547, 0.0342987980936911,
                                0.369662097773613, 0.591444264226524, 0.401225715100565, 0.666822699314234,
                                1.10327781188617, 0.827519268441487, 0.873411278211136, 1.09482906183464};
  static const double cy[20] = {-0.361066625441127, 0.0719520938231629, 0.354660608759432, 1.14336087543147,
                                -1.12006635013317, -0.457064031294315, 0.364929876813001, 0.802261227250738,
                                -0.394084977497864, -0.0315722590552472, 0.322725272389453, 1.03429879809369,
                                -1.13033790222639, 0.0137164199764117, 0.401225715100565, 1.16682269931423,
                                -0.345187707108795, -0.172480731558513, 0.373411278211136, 1.09482906183464};
  double q = 0.0, a = 0.0013;
  for (int i = 0; i < 20; ++i) {
    double R2 = (x - cx[i]) * (x - cx[i]) + (y - cy[i]) * (y - cy[i]);
    q += 1.5 * exp(-R2 / a);
  }
  return q;
}
static double q_flat(double x, double y) {  // LS_functions.hpp:54-57
  double R2 = x * x + y * y;
  return 0.5 * erfc(5.0 * (R2 - 1.0));
}
static double q_cavity(double x, double y) {  // LS_functions.hpp:59-65
  double R = sqrt(x * x + y * y), th = atan2(y, x);
  return (1.0 - pow(sin(0.5 * th), 500)) * exp(-2000 * (0.1 - R * R) * (0.1 - R * R));
}
static double q_lens(double x, double y) {  // LS_functions.hpp:67-72
  double R = sqrt(x * x + y * y);
  return 4 * (y - 0.1) * (1.0 - erf(25.0 * (R - 0.3)));
}
static double q_plasma(double x, double y) {  // LS_functions.hpp:74-93
  double C = 0.4987;
  double psi = 1.0 - (x - 0.15 * (1 - x * x)) * (x - 0.15 * (1 - x * x)) - C * (1.0 + 0.3 * x) * (1.0 + 0.3 * x) * y * y;
  double a[5] = {0.45, 0.195, 0.51, 0.195, 0.63};
  double xc[5] = {0.8, 0.54, -0.14, -0.5, 0.18};
  double yc[5] = {0.0, -0.28, 0.7, -0.01, 0.8};
  if (psi <= 0.05) return 0.0;
  double g = 0.0;
  for (int i = 0; i < 5; ++i) {
    double t = (x - xc[i]) * (x - xc[i]) + (y - yc[i]) * (y - yc[i]);
    g += a[i] * exp(-t / 0.01);
  }
  return -1.5 * (psi - 0.05) - g * cos(0.9 * y);
}

double contrast(const Params& P, double x, double y) {  // LS_functions.hpp:99-112
  switch (P.qchoice) {
    case 0: return q_gaussian(x, y);
    case 1: return q_multi(x, y);
    case 2: return q_plasma(x, y);
    case 3: return q_flat(x, y);
    case 4: return q_cavity(x, y);
    case 5: return q_lens(x, y);
    default: {
      double q = 0;
      for (size_t b = 0; b + 3 < P.bumps.size(); b += 4) {
        double a = P.bumps[b], w = P.bumps[b + 1], cx = P.bumps[b + 2], cy = P.bumps[b + 3];
        double R2 = (x - cx) * (x - cx) + (y - cy) * (y - cy);
        q += a * exp(-R2 / (w * w));
      }
      return q;
    }
  }
}

// ---------------------------------------------------------- quadrature
// Adaptive 10x10 Gauss–Legendre quadrature over [-1,1]^2
// (LS/domain2D.hpp:38-138): split a panel into 4 children (0=(-,-),
// 1=(+,-), 2=(+,+), 3=(-,+)); accept the PARENT value when
// |I_parent - sum I_children| < eps (absolute); the integral is the sum of
// accepted panels in depth-first order.
struct Quad {
  std::vector<double> t, w;
  Quad() { gauss_legendre(10, t, w); }
  template <class F>
  cd panel(F& f, double xc, double yc, double xr, double yr) const {
    cd s = 0;
    for (int m = 0; m < 10; ++m)
      for (int n = 0; n < 10; ++n) s += (w[m] * w[n]) * f(xc + xr * t[m], yc + yr * t[n]);
    return s * (xr * yr);
  }
  template <class F>
  void rec(F& f, double xc, double yc, double xr, double yr, cd Ip, double eps, cd& total) const {
    double hx = 0.5 * xr, hy = 0.5 * yr;
    double cxs[4] = {xc - hx, xc + hx, xc + hx, xc - hx};
    double cys[4] = {yc - hy, yc - hy, yc + hy, yc + hy};
    cd Ic[4];
    for (int l = 0; l < 4; ++l) Ic[l] = panel(f, cxs[l], cys[l], hx, hy);
    cd err = Ip;
    for (int l = 0; l < 4; ++l) err -= Ic[l];
    if (std::abs(err) < eps) {
      total += Ip;
    } else {
      for (int l = 0; l < 4; ++l) rec(f, cxs[l], cys[l], hx, hy, Ic[l], eps, total);
    }
  }
  template <class F>
  cd integrate(F& f, double eps) const {
    cd I0 = panel(f, 0.0, 0.0, 1.0, 1.0);
    cd total = 0;
    rec(f, 0.0, 0.0, 1.0, 1.0, I0, eps, total);
    return total;
  }
};

static double cheb_T(double x, int n) { return cos(n * acos(x)); }  // FarFieldInteraction.hpp:96-98

static double besselJn(int n, double x) {  // LS_Assembly.hpp:39-50
  if (n >= 0) return std::cyl_bessel_j((double)n, x);
  double t = std::cyl_bessel_j((double)(-n), x);
  return ((-n) % 2 == 0) ? t : -t;
}

// Round each component to 6 significant digits: the text round trip of the
// precomputation cache (LS_Assembly.hpp:1166-1228, 1383-1459; Eigen prints
// with the stream default precision 6).
static double round6(double v) {
  char buf[64];
  snprintf(buf, sizeof buf, "%.6g", v);
  return strtod(buf, nullptr);
}
static cd round6(cd v) { return cd(round6(v.real()), round6(v.imag())); }

// Source-box centres of the neighbour families in units of the target box
// radius, target box at the origin (LS_Assembly.hpp:73-120): N_centers
// (colleague), FN_centers (fine), SFN_centers (separated fine), CN_centers
// (coarse).  Source level = target level + FAM_DLEVEL (LS_Assembly.hpp:1334-1367).
static const double NC[9][2] = {{-2, -2}, {0, -2}, {2, -2}, {2, 0}, {2, 2}, {0, 2}, {-2, 2}, {-2, 0}, {0, 0}};
static const double FNC[12][2] = {{-1.5, -1.5}, {-0.5, -1.5}, {0.5, -1.5}, {1.5, -1.5}, {1.5, -0.5}, {1.5, 0.5},
                                  {1.5, 1.5},   {0.5, 1.5},   {-0.5, 1.5}, {-1.5, 1.5}, {-1.5, 0.5}, {-1.5, -0.5}};
static const double SFNC[20][2] = {{-2.5, -2.5}, {-1.5, -2.5}, {-0.5, -2.5}, {0.5, -2.5}, {1.5, -2.5},
                                   {2.5, -2.5},  {2.5, -1.5},  {2.5, -0.5},  {2.5, 0.5},  {2.5, 1.5},
                                   {2.5, 2.5},   {1.5, 2.5},   {0.5, 2.5},   {-0.5, 2.5}, {-1.5, 2.5},
                                   {-2.5, 2.5},  {-2.5, 1.5},  {-2.5, 0.5},  {-2.5, -0.5}, {-2.5, -1.5}};
static const double CNC[12][2] = {{-3, -3}, {-1, -3}, {1, -3}, {3, -3}, {3, -1}, {3, 1},
                                  {3, 3},   {1, 3},   {-1, 3}, {-3, 3}, {-3, 1}, {-3, -1}};
static const int FAM_SIZE[4] = {9, 12, 20, 12};
static const int FAM_DLEVEL[4] = {0, 1, 1, -1};
static const double (*const FAM_C[4])[2] = {NC, FNC, SFNC, CNC};

// boxRadius[j] = L / 2^j (FarFieldInteraction.hpp:42-45, LS_Assembly.hpp:886-889)
double Disc::box_radius(int j) const {
  double r = P.L;
  for (int i = 0; i < j; ++i) r = r / 2.0;
  return r;
}

// One precomputation table, unrounded.
//   family -1: far M_j = C_j Q^+, C_j(n, l) = r_j^2 int J_n(kappa r_j |s|)
//              e^{-i n th_s} T_l(s) ds  (FarFieldInteraction.hpp:60-74, 115-135, 153-160)
//   family 0..3, neighbour n: T = (V i r_B^2 / 4) Q^+, V(p, l) = int H_0(kappa
//              r_B |X_p - s|) T_l(s) ds with X_p = (r_A xi_p + 0 - c_B) / r_B and
//              c_B = centre * r_A (NearFieldInteraction1.hpp:98-128,
//              LS_Assembly.hpp:1290-1317, 1328-1367)
Mat Disc::table(int family, int level, int n) const {
  Quad quad;
  const double eps = 1e-7;  // FarFieldInteraction.hpp:37, NearFieldInteraction1.hpp:37,112
  if (family < 0) {
    const double rj = box_radius(level);
    const double kr = P.kappa * rj;
    Mat C(2 * series + 1, nb_basis);
#pragma omp parallel for schedule(dynamic) collapse(2)
    for (int m = -series; m <= series; ++m)
      for (int l = 0; l < nb_basis; ++l) {
        int a = basis_a[l], b = basis_b[l];
        auto f = [&](double x, double y) -> cd {
          double R2 = x * x + y * y;
          double normX = sqrt(R2);
          double arg = kr * normX;
          double th = atan2(y, x);
          return besselJn(m, arg) * std::exp(cd(0.0, -(double)m * th)) * cheb_T(x, a) * cheb_T(y, b);
        };
        cd I = quad.integrate(f, eps);
        C(m + series, l) = (rj * rj) * I;
      }
    return mul(C, Qpinv);
  }
  if (family > 3 || n < 0 || n >= FAM_SIZE[family]) throw std::runtime_error("no such neighbour table");
  const double ra = box_radius(level), rb = box_radius(level + FAM_DLEVEL[family]);
  const double cbx = FAM_C[family][n][0] * ra, cby = FAM_C[family][n][1] * ra;
  const double kr = P.kappa * rb;
  Mat V(rank, nb_basis);
#pragma omp parallel for schedule(dynamic) collapse(2)
  for (int p = 0; p < rank; ++p)
    for (int l = 0; l < nb_basis; ++l) {
      double Xx = (ra * cheb1d[p % P.n_cheb] + 0.0 - cbx) / rb;
      double Xy = (ra * cheb1d[p / P.n_cheb] + 0.0 - cby) / rb;
      int a = basis_a[l], b = basis_b[l];
      auto f = [&](double x, double y) -> cd {
        double R2 = (Xx - x) * (Xx - x) + (Xy - y) * (Xy - y);
        double R = sqrt(R2);
        double arg = kr * R;
        return (cd(::j0(arg), 0.0) + cd(0.0, 1.0) * ::y0(arg)) * cheb_T(x, a) * cheb_T(y, b);
      };
      V(p, l) = quad.integrate(f, eps);
    }
  Mat Vt(rank, nb_basis);
  for (size_t i = 0; i < V.a.size(); ++i) Vt.a[i] = V.a[i] * cd(0.0, 1.0) * rb * rb / 4.0;
  return mul(Vt, Qpinv);
}

void Disc::build(const Params& Pin) {
  init_basis(Pin);
  // points: leaf k, index jy*n+kx at center + r*(xi_kx, xi_jy) (LS_Assembly.hpp:1074-1087, 543-550)
  px.resize(N);
  py.resize(N);
  q.resize(N);
  for (int k = 0; k < n_leaf; ++k) {
    int ix, iy;
    box_xy(P.n_levels, k, ix, iy);
    double cx = -P.L + (2 * ix + 1) * r_leaf, cy = -P.L + (2 * iy + 1) * r_leaf;
    for (int jy = 0; jy < P.n_cheb; ++jy)
      for (int kx = 0; kx < P.n_cheb; ++kx) {
        int64_t i = (int64_t)k * rank + jy * P.n_cheb + kx;
        px[i] = r_leaf * cheb1d[kx] + cx;
        py[i] = r_leaf * cheb1d[jy] + cy;
      }
  }
  for (int64_t i = 0; i < N; ++i) q[i] = contrast(P, px[i], py[i]);
  // the tables the uniform tree reads: M and the colleague operators at the
  // leaf level (LS_Assembly.hpp:496, 527)
  M = table(-1, P.n_levels, 0);
  for (int nb = 0; nb < 9; ++nb) T[nb] = table(0, P.n_levels, nb);
  if (P.quantize) {
    for (auto& v : M.a) v = round6(v);
    for (int nb = 0; nb < 9; ++nb)
      for (auto& v : T[nb].a) v = round6(v);
  }
}

cd round6c(cd v) { return round6(v); }

// Chebyshev nodes, the basis order and Q^+ (no points, no tables).
void Disc::init_basis(const Params& Pin) {
  P = Pin;
  rank = P.n_cheb * P.n_cheb;
  nb_basis = P.degree * (P.degree + 1) / 2;
  n_leaf = 1 << (2 * P.n_levels);
  N = (int64_t)n_leaf * rank;
  double rad = P.L;
  for (int j = 0; j < P.n_levels; ++j) rad *= 0.5;
  r_leaf = rad;
  // Chebyshev nodes (LS_Assembly.hpp:567-580)
  cheb1d.resize(P.n_cheb);
  for (int k = 0; k < P.n_cheb; ++k) cheb1d[k] = -cos((k + 0.5) / P.n_cheb * PI);
  // basis order (FarFieldInteraction.hpp:46-54)
  basis_a.clear();
  basis_b.clear();
  for (int d = 0; d < P.degree; ++d)
    for (int a = 0; a < d + 1; ++a) basis_a.push_back(a), basis_b.push_back(d - a);
  // Q and its pseudo-inverse (FarFieldInteraction.hpp:138-150)
  Mat Qm(rank, nb_basis);
  for (int i = 0; i < rank; ++i) {
    double sx = cheb1d[i % P.n_cheb], sy = cheb1d[i / P.n_cheb];
    for (int l = 0; l < nb_basis; ++l) Qm(i, l) = cheb_T(sx, basis_a[l]) * cheb_T(sy, basis_b[l]);
  }
  {
    // full column rank: pinv = P R^{-1} Q1^H from the pivoted QR
    CPQR f;
    f.compute(Qm);
    Mat Q1 = f.Q(nb_basis);
    Mat Rinv_Q1H(nb_basis, rank);
    Mat QH = adj(Q1);
    for (int col = 0; col < rank; ++col) {
      for (int i = nb_basis - 1; i >= 0; --i) {
        cd s = QH(i, col);
        for (int l = i + 1; l < nb_basis; ++l) s -= f.qr(i, l) * Rinv_Q1H(l, col);
        Rinv_Q1H(i, col) = s / f.qr(i, i);
      }
    }
    Qpinv = Mat(nb_basis, rank);
    for (int i = 0; i < nb_basis; ++i)
      for (int col = 0; col < rank; ++col) Qpinv(f.perm[i], col) = Rinv_Q1H(i, col);
  }
}

// Graf far-field sum for target point i and source leaf box (LS_Assembly.hpp:518-529).
cd Disc::far_sum(int64_t i, int leaf_j, int idx_j) const {
  int ix, iy;
  box_xy(P.n_levels, leaf_j, ix, iy);
  double cx = -P.L + (2 * ix + 1) * r_leaf, cy = -P.L + (2 * iy + 1) * r_leaf;
  return graf_series(M, idx_j, px[i], py[i], cx, cy, P.kappa, series, P.exact_bessel != 0);
}

// sum_{n=-series}^{series} M(n + series, idx_j) H_n(kappa rho) e^{i n phi}, with
// (rho, phi) the polar form of (x - cx, y - cy) (LS_Assembly.hpp:518-529)
cd graf_series(const Mat& M, int idx_j, double x, double y, double cx, double cy, double kappa, int series,
               bool exact) {
  double phi = atan2(y - cy, x - cx);
  double R2 = (y - cy) * (y - cy) + (x - cx) * (x - cx);
  double R = kappa * sqrt(R2);
  double J[64], Y[64];
  bessel_jy(R, series, J, Y, exact);
  cd s = 0;
  if (exact) {
    for (int n = -series; n <= series; ++n) {
      int an = n < 0 ? -n : n;
      double sg = (n < 0 && (an % 2)) ? -1.0 : 1.0;
      cd H(sg * J[an], sg * Y[an]);
      cd temp = H * std::exp(cd(0.0, 1.0) * (double)n * phi);
      s += M(n + series, idx_j) * temp;
    }
  } else {
    cd e1 = std::exp(cd(0.0, phi));
    cd pw[61];
    pw[0] = 1.0;
    for (int n = 1; n <= series; ++n) pw[n] = pw[n - 1] * e1;
    for (int n = -series; n <= series; ++n) {
      int an = n < 0 ? -n : n;
      double sg = (n < 0 && (an % 2)) ? -1.0 : 1.0;
      cd H(sg * J[an], sg * Y[an]);
      cd e = n < 0 ? std::conj(pw[an]) : pw[an];
      s += M(n + series, idx_j) * (H * e);
    }
  }
  return s;
}

// A(i,j) (LS_Assembly.hpp:474-541; kernel.hpp:130-136).
cd Disc::entry(int64_t i, int64_t j) const {
  int li = (int)(i / rank), ii = (int)(i % rank), lj = (int)(j / rank), jj = (int)(j % rank);
  int xi, yi, xj, yj;
  box_xy(P.n_levels, li, xi, yi);
  box_xy(P.n_levels, lj, xj, yj);
  int dx = xj - xi, dy = yj - yi;
  if (dx >= -1 && dx <= 1 && dy >= -1 && dy <= 1) {
    int nb;
    if (dx == 0 && dy == 0) nb = 8;
    else {
      nb = -1;
      for (int s = 0; s < 8; ++s)
        if (NB_OFF[s][0] == dx && NB_OFF[s][1] == dy) nb = s;
    }
    cd e = T[nb](ii, jj);
    if (find_phi) {  // LS_Assembly.hpp:507-512
      e = e * P.kappa;
      e = e * P.kappa;
      e = e * q[i];
      if (i == j) e = 1.0 + e;
    }
    if (std::abs(e) < 1e-16) e = 0.0;
    return e;
  }
  cd s = far_sum(i, lj, jj);
  if (perturb != 0) {
    uint64_t h = ((uint64_t)i * 0x9E3779B97F4A7C15ull) ^ ((uint64_t)j * 0xC2B2AE3D27D4EB4Full) ^ perturb_seed;
    h ^= h >> 31;
    h *= 0xBF58476D1CE4E5B9ull;
    h ^= h >> 29;
    s *= 1.0 + perturb * ((double)(h % 2000001ull) / 1000000.0 - 1.0);
  }
  s *= cd(0.0, 0.25);
  if (find_phi) {  // LS_Assembly.hpp:530-535
    s *= P.kappa * P.kappa * q[i];
    if (i == j) s += 1.0;
  }
  if (std::abs(s) < 1e-16) s = 0.0;
  return s;
}

}  // namespace orc
