// Problem set-up on the host (points, contrast, lists, task lists) and the
// device work of the assembly phase: table quadrature, NCA via batched ACA,
// L2P/P2M transfer operators, M2L and leaf P2P blocks.
#include "problem.h"

#include <algorithm>
#include <array>

#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>

namespace aifmm {

static double now_s() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

static const double PI = 3.1415926535897932384;  // LS_Assembly.hpp:30

// neighbour / interaction-list slot offsets (FMM2DTree_find_error.hpp:168-1046)
static const int NB_OFF[8][2] = {{-1, -1}, {0, -1}, {1, -1}, {1, 0}, {1, 1}, {0, 1}, {-1, 1}, {-1, 0}};
static const int IN_OFF[16][2] = {{-2, -2}, {-1, -2}, {0, -2}, {1, -2}, {2, -2}, {2, -1}, {2, 0}, {2, 1},
                                  {2, 2},   {1, 2},   {0, 2},  {-1, 2}, {-2, 2}, {-2, 1}, {-2, 0}, {-2, -1}};
static const int OUT_OFF[24][2] = {{-3, -3}, {-2, -3}, {-1, -3}, {0, -3}, {1, -3}, {2, -3}, {3, -3}, {3, -2},
                                   {3, -1},  {3, 0},   {3, 1},   {3, 2},  {3, 3},  {2, 3},  {1, 3},  {0, 3},
                                   {-1, 3},  {-2, 3},  {-3, 3},  {-3, 2}, {-3, 1}, {-3, 0}, {-3, -1}, {-3, -2}};

// Contrast functions q(x) (LS/LS_functions.hpp:1-112); 6 = sum of Gaussian
// bumps a*exp(-|x-c|^2/w^2).
double contrast(const Params& P, double x, double y) {
  switch (P.qchoice) {
    case 0: {
      double R2 = x * x + y * y;
      return 1.5 * exp(-160.0 * R2);
    }
    case 1: {
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
      double qv = 0.0;
      for (int i = 0; i < 20; ++i) {
        double R2 = (x - cx[i]) * (x - cx[i]) + (y - cy[i]) * (y - cy[i]);
        qv += 1.5 * exp(-R2 / 0.0013);
      }
      return qv;
    }
    case 2: {
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
    case 3: {
      double R2 = x * x + y * y;
      return 0.5 * erfc(5.0 * (R2 - 1.0));
    }
    case 4: {
      double R = sqrt(x * x + y * y), th = atan2(y, x);
      return (1.0 - pow(sin(0.5 * th), 500)) * exp(-2000 * (0.1 - R * R) * (0.1 - R * R));
    }
    case 5: {
      double R = sqrt(x * x + y * y);
      return 4 * (y - 0.1) * (1.0 - erf(25.0 * (R - 0.3)));
    }
    default: {
      double qv = 0;
      for (size_t b = 0; b + 3 < P.bumps.size(); b += 4) {
        double a = P.bumps[b], w = P.bumps[b + 1], cx = P.bumps[b + 2], cy = P.bumps[b + 3];
        double R2 = (x - cx) * (x - cx) + (y - cy) * (y - cy);
        qv += a * exp(-R2 / (w * w));
      }
      return qv;
    }
  }
}

// 10-point Gauss–Legendre rule by Newton iteration on P_10 in long double,
// ordered (-x, +x) with |x| ascending.
void gauss_legendre10(std::vector<double>& x, std::vector<double>& w) {
  const int n = 10;
  x.clear();
  w.clear();
  std::vector<long double> pos, wp;
  for (int i = 0; i < n / 2; ++i) {
    long double z = cosl(3.14159265358979323846264338327950288L * (i + 0.75L) / (n + 0.5L)), dp = 0;
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
    wp.push_back(2 / ((1 - z * z) * dp * dp));
  }
  for (int i = n / 2 - 1; i >= 0; --i) {
    x.push_back((double)-pos[i]);
    w.push_back((double)wp[i]);
    x.push_back((double)pos[i]);
    w.push_back((double)wp[i]);
  }
}

static double round6(double v) {
  char buf[64];
  snprintf(buf, sizeof buf, "%.6g", v);
  return strtod(buf, nullptr);
}

Problem::Problem(const Params& Pin, const std::string& precomp_dir) : P(Pin) {
  // non-blocking: no implicit ordering with the null stream, so problems on
  // different host threads overlap on the device
  stream = stream_create(hipStreamNonBlocking);
  try {
    init(precomp_dir);
  } catch (...) {
    stream_destroy(stream);
    stream = nullptr;
    throw;
  }
}

void Problem::init(const std::string& precomp_dir) {
  R = P.n_cheb * P.n_cheb;
  nbasis = P.degree * (P.degree + 1) / 2;
  if (P.adaptive) {
    if (!precomp_dir.empty()) throw std::runtime_error("adaptive discretisation: tables are computed, not read");
    init_adaptive();
    return;
  }
  L = P.n_levels;
  if (L < 2 || L > 9) throw std::runtime_error("n_levels must be in [2, 9]");
  nleaf = 1 << (2 * L);
  N = (int64_t)nleaf * R;
  // arena chunks sized by N, so that small problems stay small (a cache of
  // resident problems is admitted by their device bytes)
  persist.set_chunk(Arena::chunk_for(N, 64, 4 << 20, 64 << 20));
  op.set_chunk(Arena::chunk_for(N, 8 << 10, 32 << 20, 1 << 30));
  acaws.set_chunk(Arena::chunk_for(N, 8 << 10, 32 << 20, 1 << 30));
  scratch.set_chunk(Arena::chunk_for(N, 2 << 10, 16 << 20, 256 << 20));
  acakeep.set_chunk(Arena::chunk_for(N, 2 << 10, 16 << 20, 256 << 20));
  rleaf = P.L;
  for (int j = 0; j < L; ++j) rleaf *= 0.5;
  // points: leaf k holds points 36k..36k+35, index jy*n+kx at
  // c_k + r (xi_kx, xi_jy) (LS_Assembly.hpp:543-607, 1074-1087)
  std::vector<double> cheb(P.n_cheb);
  for (int k = 0; k < P.n_cheb; ++k) cheb[k] = -cos((k + 0.5) / P.n_cheb * PI);
  px.resize(N);
  py.resize(N);
  q.resize(N);
  for (int k = 0; k < nleaf; ++k) {
    int ix, iy;
    box_xy(L, k, ix, iy);
    double cx = -P.L + (2 * ix + 1) * rleaf, cy = -P.L + (2 * iy + 1) * rleaf;
    for (int jy = 0; jy < P.n_cheb; ++jy)
      for (int kx = 0; kx < P.n_cheb; ++kx) {
        int64_t i = (int64_t)k * R + jy * P.n_cheb + kx;
        px[i] = rleaf * cheb[kx] + cx;
        py[i] = rleaf * cheb[jy] + cy;
      }
  }
  for (int64_t i = 0; i < N; ++i) q[i] = contrast(P, px[i], py[i]);
  dpx = persist.get<double>(N);
  dpy = persist.get<double>(N);
  dq = persist.get<double>(N);
  h2d(dpx, px.data(), N * 8, stream);
  h2d(dpy, py.data(), N * 8, stream);
  h2d(dq, q.data(), N * 8, stream);
  std::vector<int64_t> iota(N);
  for (int64_t i = 0; i < N; ++i) iota[i] = i;
  d_iota = persist.get<int64_t>(N);
  h2d(d_iota, iota.data(), N * 8, stream);
  build_lists();
  if (precomp_dir.empty())
    build_tables();
  else
    load_tables(precomp_dir);
}

Problem::~Problem() {
  if (!stream) return;
  stream_destroy(stream);  // synchronises first: queued work still reads this problem's memory
}

void Problem::build_lists() {
  lists.assign(L + 1, {});
  for (int j = 0; j <= L; ++j) {
    int nb = 1 << (2 * j), side = 1 << j;
    lists[j].resize(nb);
    for (int k = 0; k < nb; ++k) {
      Lists& B = lists[j][k];
      for (int s = 0; s < 8; ++s) B.nb[s] = -1;
      for (int s = 0; s < 16; ++s) B.inner[s] = -1;
      for (int s = 0; s < 24; ++s) B.outer[s] = -1;
      if (j == 0) continue;
      int ix, iy;
      box_xy(j, k, ix, iy);
      auto in = [&](int x, int y) { return x >= 0 && y >= 0 && x < side && y < side; };
      for (int s = 0; s < 8; ++s) {
        int x = ix + NB_OFF[s][0], y = iy + NB_OFF[s][1];
        if (in(x, y)) B.nb[s] = xy_box(j, x, y);
      }
      for (int s = 0; s < 16; ++s) {
        int x = ix + IN_OFF[s][0], y = iy + IN_OFF[s][1];
        if (in(x, y)) B.inner[s] = xy_box(j, x, y);
      }
      for (int s = 0; s < 24; ++s) {
        int x = ix + OUT_OFF[s][0], y = iy + OUT_OFF[s][1];
        if (in(x, y) && std::abs((x >> 1) - (ix >> 1)) <= 1 && std::abs((y >> 1) - (iy >> 1)) <= 1)
          B.outer[s] = xy_box(j, x, y);
      }
    }
  }
}

int Problem::well_sep(int j, int k, int l) const {  // FMM2DTree_find_error.hpp:2077-2088
  if (k == l) return 0;
  for (int i = 0; i < 8; ++i)
    if (lists[j][k].nb[i] == l) return 0;
  return 1;
}

// Pseudo-inverse of the 36x21 Chebyshev collocation matrix Q
// (FarFieldInteraction.hpp:138-150) by modified Gram–Schmidt in long double.
static std::vector<double> q_pinv(int n_cheb, const std::vector<int>& ba, const std::vector<int>& bb) {
  int R = n_cheb * n_cheb, nb = (int)ba.size();
  std::vector<long double> Q((size_t)R * nb), Rm((size_t)nb * nb, 0.0L);
  std::vector<long double> cheb(n_cheb);
  for (int k = 0; k < n_cheb; ++k) cheb[k] = -cosl((k + 0.5L) / n_cheb * 3.14159265358979323846264338327950288L);
  for (int i = 0; i < R; ++i)
    for (int l = 0; l < nb; ++l)
      Q[i + (size_t)R * l] = cosl(ba[l] * acosl(cheb[i % n_cheb])) * cosl(bb[l] * acosl(cheb[i / n_cheb]));
  for (int l = 0; l < nb; ++l) {
    for (int pass = 0; pass < 2; ++pass)
      for (int k = 0; k < l; ++k) {
        long double d = 0;
        for (int i = 0; i < R; ++i) d += Q[i + (size_t)R * k] * Q[i + (size_t)R * l];
        Rm[k + (size_t)nb * l] += d;
        for (int i = 0; i < R; ++i) Q[i + (size_t)R * l] -= d * Q[i + (size_t)R * k];
      }
    long double nr = 0;
    for (int i = 0; i < R; ++i) nr += Q[i + (size_t)R * l] * Q[i + (size_t)R * l];
    nr = sqrtl(nr);
    Rm[l + (size_t)nb * l] = nr;
    for (int i = 0; i < R; ++i) Q[i + (size_t)R * l] /= nr;
  }
  // pinv = R^{-1} Q^T  (nb x R)
  std::vector<double> out((size_t)nb * R);
  for (int c = 0; c < R; ++c) {
    std::vector<long double> y(nb);
    for (int i = nb - 1; i >= 0; --i) {
      long double s = Q[c + (size_t)R * i];
      for (int k = i + 1; k < nb; ++k) s -= Rm[i + (size_t)nb * k] * y[k];
      y[i] = s / Rm[i + (size_t)nb * i];
    }
    for (int i = 0; i < nb; ++i) out[i + (size_t)nb * c] = (double)y[i];
  }
  return out;
}

// Source-box centres of the four neighbour families, in units of the target
// box radius r_A, target box at the origin (LS_Assembly.hpp:73-120): colleague
// (N_centers), fine (FN_centers), separated fine (SFN_centers) and coarse
// (CN_centers) neighbours.  The source box level is level_A + dlevel
// (LS_Assembly.hpp:1334-1367).
static const double FAM_C0[9][2] = {{-2, -2}, {0, -2}, {2, -2}, {2, 0}, {2, 2}, {0, 2}, {-2, 2}, {-2, 0}, {0, 0}};
static const double FAM_C1[12][2] = {{-1.5, -1.5}, {-0.5, -1.5}, {0.5, -1.5}, {1.5, -1.5}, {1.5, -0.5}, {1.5, 0.5},
                                     {1.5, 1.5},   {0.5, 1.5},   {-0.5, 1.5}, {-1.5, 1.5}, {-1.5, 0.5}, {-1.5, -0.5}};
static const double FAM_C2[20][2] = {{-2.5, -2.5}, {-1.5, -2.5}, {-0.5, -2.5}, {0.5, -2.5}, {1.5, -2.5},
                                     {2.5, -2.5},  {2.5, -1.5},  {2.5, -0.5},  {2.5, 0.5},  {2.5, 1.5},
                                     {2.5, 2.5},   {1.5, 2.5},   {0.5, 2.5},   {-0.5, 2.5}, {-1.5, 2.5},
                                     {-2.5, 2.5},  {-2.5, 1.5},  {-2.5, 0.5},  {-2.5, -0.5}, {-2.5, -1.5}};
static const double FAM_C3[12][2] = {{-3, -3}, {-1, -3}, {1, -3}, {3, -3}, {3, -1}, {3, 1},
                                     {3, 3},   {1, 3},   {-1, 3}, {-3, 3}, {-3, 1}, {-3, -1}};
const int FAMILY_SIZE[4] = {9, 12, 20, 12};
static const int FAMILY_DLEVEL[4] = {0, 1, 1, -1};
static const double (*const FAMILY_CENTRES[4])[2] = {FAM_C0, FAM_C1, FAM_C2, FAM_C3};

double Problem::box_radius(int j) const {  // boxRadius[j] (LS_Assembly.hpp:886-889)
  double r = P.L;
  for (int i = 0; i < j; ++i) r = r / 2.0;
  return r;
}

// The tables named by specs, unrounded, as computed by the reference's
// precomputation (FarFieldInteraction.hpp:115-160; LS_Assembly.hpp:1290-1317):
//   far  M_j = (r_j^2 C_j) Q^+                       NTERM x R
//   near T   = (i r_B^2 / 4) V Q^+                   R x R
// The integrals run on the device, one wave per integral; Q^+ and the
// products follow on the device.  Column-major results, one per spec.
std::vector<std::vector<double2>> Problem::quad_tables(const std::vector<TableSpec>& specs, int* capped) {
  std::vector<int> ba, bb;
  for (int d = 0; d < P.degree; ++d)
    for (int a = 0; a < d + 1; ++a) ba.push_back(a), bb.push_back(d - a);
  std::vector<double> cheb(P.n_cheb), glx, glw;
  for (int k = 0; k < P.n_cheb; ++k) cheb[k] = -cos((k + 0.5) / P.n_cheb * PI);
  gauss_legendre10(glx, glw);
  std::vector<QuadTask> tasks;
  std::vector<QuadGeom> geo;
  std::vector<size_t> first(specs.size());
  for (size_t s = 0; s < specs.size(); ++s) {
    const TableSpec& t = specs[s];
    first[s] = tasks.size();
    int g = (int)geo.size();
    if (t.family < 0) {
      if (t.level < 0) throw std::runtime_error("far table level < 0");
      geo.push_back(QuadGeom{P.kappa * box_radius(t.level), 0, 0, 0, 0});
      for (int n = -SERIES; n <= SERIES; ++n)
        for (int l = 0; l < nbasis; ++l) tasks.push_back(QuadTask{0, n, 0, l, g});
    } else {
      if (t.family > 3 || t.n < 0 || t.n >= FAMILY_SIZE[t.family]) throw std::runtime_error("no such neighbour table");
      int lb = t.level + FAMILY_DLEVEL[t.family];
      if (t.level < 0 || lb < 0) throw std::runtime_error("neighbour table level out of range");
      double ra = box_radius(t.level), rb = box_radius(lb);
      const double* c = FAMILY_CENTRES[t.family][t.n];
      geo.push_back(QuadGeom{P.kappa * rb, ra, rb, c[0] * ra, c[1] * ra});
      for (int p = 0; p < R; ++p)
        for (int l = 0; l < nbasis; ++l) tasks.push_back(QuadTask{1, 0, p, l, g});
    }
  }
  QuadCtx Q;
  Q.basis_a = upload(scratch, ba, stream);
  Q.basis_b = upload(scratch, bb, stream);
  Q.cheb1d = upload(scratch, cheb, stream);
  Q.gl_t = upload(scratch, glx, stream);
  Q.gl_w = upload(scratch, glw, stream);
  Q.geom = upload(scratch, geo, stream);
  Q.n_cheb = P.n_cheb;
  Q.eps = 1e-7;  // FarFieldInteraction.hpp:37, NearFieldInteraction1.hpp:37, 112
  QuadTask* dt = upload(scratch, tasks, stream);
  double2* dint = scratch.get<double2>(tasks.size());
  int* dflag = scratch.get<int>(1);
  HIP_OK(hipMemsetAsync(dflag, 0, 4, stream));
  launch_quadrature(Q, (int)tasks.size(), dt, dint, dflag, stream);
  std::vector<double> qp = q_pinv(P.n_cheb, ba, bb);
  std::vector<double2> qpc(qp.size());
  for (size_t i = 0; i < qp.size(); ++i) qpc[i] = make_double2(qp[i], 0.0);
  Blk<double2> QP{upload(scratch, qpc, stream), nbasis, R, nbasis};
  // far tasks are laid out [n][l]: column-major C^T (nbasis x NTERM); near
  // tasks [p][l]: V^T (nbasis x R).  Transpose, then multiply by Q^+.
  CopyBatch<double2> cp;
  GemmBatch<double2> g;
  std::vector<Blk<double2>> outb(specs.size());
  for (size_t s = 0; s < specs.size(); ++s) {
    bool far = specs[s].family < 0;
    int rows = far ? NTERM : R;
    Blk<double2> XT{dint + first[s], nbasis, rows, nbasis};
    Blk<double2> X = new_blk<double2>(scratch, rows, nbasis);
    cp.add(XT, X, 2);
    outb[s] = new_blk<double2>(scratch, rows, R);
    double r;
    if (far) {
      r = box_radius(specs[s].level);
      g.add(X, 0, QP, 0, outb[s], make_double2(r * r, 0), make_double2(0, 0));
    } else {
      r = box_radius(specs[s].level + FAMILY_DLEVEL[specs[s].family]);
      g.add(X, 0, QP, 0, outb[s], make_double2(0, r * r / 4.0), make_double2(0, 0));
    }
  }
  cp.run(scratch, stream);
  g.run(scratch, stream);
  std::vector<std::vector<double2>> out(specs.size());
  for (size_t s = 0; s < specs.size(); ++s) {
    out[s].resize(outb[s].size());
    d2h(out[s].data(), outb[s].p, out[s].size() * 16, stream);
  }
  int cap = 0;
  d2h(&cap, dflag, 4, stream);
  HIP_OK(hipStreamSynchronize(stream));
  scratch.reset();
  if (capped) *capped = cap;
  return out;
}

// Far-field M and colleague operators T_nb at the leaf level, the tables the
// uniform tree reads (LS_Assembly.hpp:496, 527).  With quantize, the
// 6-significant-digit rounding of the reference's text cache
// (LS_Assembly.hpp:1166-1228, 1383-1459) is applied to the finished tables.
void Problem::build_tables() {
  double t0 = now_s();
  std::vector<TableSpec> specs{{-1, L, 0}};
  for (int nb = 0; nb < 9; ++nb) specs.push_back({0, L, nb});
  auto tabs = quad_tables(specs, &quad_capped);
  M_h = std::move(tabs[0]);
  T_h.clear();
  for (int nb = 0; nb < 9; ++nb) T_h.insert(T_h.end(), tabs[1 + nb].begin(), tabs[1 + nb].end());
  if (P.quantize) {
    for (auto& v : M_h) v = make_double2(round6(v.x), round6(v.y));
    for (auto& v : T_h) v = make_double2(round6(v.x), round6(v.y));
  }
  install_tables();
  t_tables = now_s() - t0;
}

// M_h and T_h to the device and into the entry context.
void Problem::install_tables() {
  dM = persist.get<double2>(M_h.size());
  dT = persist.get<double2>(T_h.size());
  h2d(dM, M_h.data(), M_h.size() * 16, stream);
  h2d(dT, T_h.data(), T_h.size() * 16, stream);
  ctx.px = dpx;
  ctx.py = dpy;
  ctx.q = dq;
  ctx.M = dM;
  ctx.T = dT;
  ctx.R = R;
  ctx.L = L;
  ctx.kappa = P.kappa;
  ctx.rleaf = rleaf;
  ctx.Ldom = P.L;
}

void Problem::need_uniform(const char* what) const {
  if (P.adaptive)
    throw std::runtime_error(std::string(what) +
                             ": needs the uniform tree (the reference's FMM2DTree solver works on it); this problem "
                             "has the adaptive discretisation");
}

// The adaptive discretisation (Params::adaptive): the leaves of the adaptive
// quadtree with 2:1 level restriction in the reference's leafNodes order,
// points per leaf as assignLeafChargeLocations places them
// (LS_Assembly.hpp:1074-1087), each leaf's near leaves with the neighbour
// table getIfNeighbor picks, and the tables: the far M at every leaf level
// and every near table a pair uses, by device quadrature, rounded to 6
// digits with quantize as the text cache would.
void Problem::init_adaptive() {
  double t0 = now_s();
  leaves = adaptive_leaves(P, P.tree_adaptivity, &L);
  nleaf = (int)leaves.size();
  N = (int64_t)nleaf * R;
  persist.set_chunk(Arena::chunk_for(N, 64, 4 << 20, 64 << 20));
  scratch.set_chunk(Arena::chunk_for(N, 2 << 10, 16 << 20, 256 << 20));
  std::vector<double> cheb(P.n_cheb);
  for (int k = 0; k < P.n_cheb; ++k) cheb[k] = -cos((k + 0.5) / P.n_cheb * PI);
  px.resize(N);
  py.resize(N);
  q.resize(N);
  std::vector<double> rad(L + 1);
  for (int j = 0; j <= L; ++j) rad[j] = box_radius(j);
  rleaf = rad[L];
  for (int k = 0; k < nleaf; ++k) {
    const AdaptiveLeaf& f = leaves[k];
    const double r = rad[f.level];
    for (int jy = 0; jy < P.n_cheb; ++jy)
      for (int kx = 0; kx < P.n_cheb; ++kx) {
        int64_t i = (int64_t)k * R + jy * P.n_cheb + kx;
        px[i] = r * cheb[kx] + f.cx;
        py[i] = r * cheb[jy] + f.cy;
      }
  }
  for (int64_t i = 0; i < N; ++i) q[i] = contrast(P, px[i], py[i]);
  // near leaves: candidates from a grid of cells of the coarsest leaf size
  int lmin = L;
  for (auto& f : leaves) lmin = std::min(lmin, f.level);
  const int side = 1 << lmin;
  const double cell = 2.0 * rad[lmin];
  auto cell_of = [&](double v) { return std::min(side - 1, std::max(0, (int)floor((v + P.L) / cell))); };
  std::vector<std::vector<int>> grid((size_t)side * side);
  for (int k = 0; k < nleaf; ++k) grid[(size_t)cell_of(leaves[k].cy) * side + cell_of(leaves[k].cx)].push_back(k);
  std::map<std::array<int, 3>, int> tab_of;
  near_off.assign(nleaf + 1, 0);
  near_leaf.clear();
  near_tab.clear();
  near_specs.clear();
  for (int a = 0; a < nleaf; ++a) {
    const AdaptiveLeaf& fa = leaves[a];
    const double w = 3.0 * rad[fa.level] + rad[lmin];
    const int x0 = cell_of(fa.cx - w), x1 = cell_of(fa.cx + w), y0 = cell_of(fa.cy - w), y1 = cell_of(fa.cy + w);
    std::vector<int> cand;
    for (int cy = y0; cy <= y1; ++cy)
      for (int cx = x0; cx <= x1; ++cx)
        for (int b : grid[(size_t)cy * side + cx]) cand.push_back(b);
    std::sort(cand.begin(), cand.end());
    for (int b : cand) {
      int slot = -1;
      const int fam = classify_near(fa, leaves[b], rad[fa.level], rad[leaves[b].level], slot);
      if (fam < 0) continue;
      const std::array<int, 3> key{fam, fa.level, slot};
      auto it = tab_of.find(key);
      if (it == tab_of.end()) {
        it = tab_of.emplace(key, (int)near_specs.size()).first;
        near_specs.push_back(TableSpec{fam, fa.level, slot});
      }
      near_leaf.push_back(b);
      near_tab.push_back(it->second);
    }
    near_off[a + 1] = (int)near_leaf.size();
  }
  // tables: the far M of every level 0..L (levels without leaves are never read)
  std::vector<TableSpec> specs;
  for (int j = 0; j <= L; ++j) specs.push_back({-1, j, 0});
  specs.insert(specs.end(), near_specs.begin(), near_specs.end());
  auto tabs = quad_tables(specs, &quad_capped);
  if (P.quantize)
    for (auto& t : tabs)
      for (auto& v : t) v = make_double2(round6(v.x), round6(v.y));
  std::vector<double2> Mlev, Ttab;
  for (int j = 0; j <= L; ++j) Mlev.insert(Mlev.end(), tabs[j].begin(), tabs[j].end());
  for (size_t t = 0; t < near_specs.size(); ++t) Ttab.insert(Ttab.end(), tabs[L + 1 + t].begin(), tabs[L + 1 + t].end());
  M_h = tabs[L];
  std::vector<int> lv(nleaf);
  std::vector<double> lcx(nleaf), lcy(nleaf);
  for (int k = 0; k < nleaf; ++k) lv[k] = leaves[k].level, lcx[k] = leaves[k].cx, lcy[k] = leaves[k].cy;
  dpx = upload(persist, px, stream);
  dpy = upload(persist, py, stream);
  dq = upload(persist, q, stream);
  ctx = EntryCtx{};
  ctx.px = dpx;
  ctx.py = dpy;
  ctx.q = dq;
  ctx.R = R;
  ctx.L = L;
  ctx.kappa = P.kappa;
  ctx.rleaf = rleaf;
  ctx.Ldom = P.L;
  ctx.adaptive = 1;
  ctx.leaf_lv = upload(persist, lv, stream);
  ctx.lcx = upload(persist, lcx, stream);
  ctx.lcy = upload(persist, lcy, stream);
  ctx.rad = upload(persist, rad, stream);
  ctx.Mlev = upload(persist, Mlev, stream);
  ctx.near_off = upload(persist, near_off, stream);
  ctx.near_leaf = near_leaf.empty() ? nullptr : upload(persist, near_leaf, stream);
  ctx.near_tab = near_tab.empty() ? nullptr : upload(persist, near_tab, stream);
  ctx.Ttab = Ttab.empty() ? nullptr : upload(persist, Ttab, stream);
  HIP_OK(hipStreamSynchronize(stream));
  t_tables = now_s() - t0;
}

void Problem::entries(int64_t n, const int64_t* I, const int64_t* J, double2* out) {
  std::vector<int64_t> Iv(I, I + n), Jv(J, J + n);
  int64_t* dI = upload(scratch, Iv, stream);
  int64_t* dJ = upload(scratch, Jv, stream);
  double2* dout = scratch.get<double2>(n);
  launch_entries(ctx_phi(), n, dI, dJ, dout, stream);
  d2h(out, dout, n * 16, stream);
  HIP_OK(hipStreamSynchronize(stream));
  scratch.reset();
}

// ------------------------------------------------------------------ NCA
// getNodes_outgoing_box / getNodes_incoming_box / get_L2P_P2M_box
// (FMM2DTree_find_error.hpp:1105-1336) for every box of level j at once.
void Problem::nca_level(int j) {
  int nb = 1 << (2 * j);
  auto& S = sk[j];
  // index lists: per box 4 lists (out rows=IL, out cols=own, in rows=own, in cols=IL)
  std::vector<int64_t> idx;
  std::vector<size_t> off(4 * nb + 1);
  auto own_in = [&](int k, std::vector<int64_t>& v) {  // incoming check candidates
    if (j == L)
      for (int i = 0; i < R; ++i) v.push_back((int64_t)k * R + i);
    else
      for (int c = 0; c < 4; ++c) v.insert(v.end(), sk[j + 1][4 * k + c].in_chk.begin(), sk[j + 1][4 * k + c].in_chk.end());
  };
  auto own_out = [&](int k, std::vector<int64_t>& v) {  // outgoing charge candidates
    if (j == L)
      for (int i = 0; i < R; ++i) v.push_back((int64_t)k * R + i);
    else
      for (int c = 0; c < 4; ++c)
        v.insert(v.end(), sk[j + 1][4 * k + c].out_chg.begin(), sk[j + 1][4 * k + c].out_chg.end());
  };
  std::vector<std::vector<int64_t>> L4(4 * nb);
  for (int k = 0; k < nb; ++k) {
    const Lists& B = lists[j][k];
    auto& outR = L4[4 * k + 0];
    auto& outC = L4[4 * k + 1];
    auto& inR = L4[4 * k + 2];
    auto& inC = L4[4 * k + 3];
    own_out(k, outC);
    own_in(k, inR);
    for (int s = 0; s < 16; ++s)
      if (B.inner[s] != -1) own_in(B.inner[s], outR), own_out(B.inner[s], inC);
    for (int s = 0; s < 24; ++s)
      if (B.outer[s] != -1) own_in(B.outer[s], outR), own_out(B.outer[s], inC);
  }
  for (int t = 0; t < 4 * nb; ++t) {
    off[t] = idx.size();
    idx.insert(idx.end(), L4[t].begin(), L4[t].end());
  }
  off[4 * nb] = idx.size();
  int64_t* didx = upload(scratch, idx, stream);
  // ACA problems, chunked by workspace size
  struct Res {
    int rank = 0;
    std::vector<int> pb, sb;
    Blk<double2> keep;
    int np = 0, ns = 0, prim_rows = 0;
  };
  std::vector<Res> res(2 * nb);
  const double tol = pow(10, -1.0 * P.tol_pow), tol_robust = 1e-12;
  const size_t budget = size_t(4) << 30;
  int p0 = 0;
  while (p0 < 2 * nb) {
    std::vector<AcaProblem> probs;
    std::vector<size_t> roff;  // problem q: rank at roff[q], then rm primary and rm secondary pivots
    size_t rtot = 0;
    size_t bytes = 0;
    int p1 = p0;
    for (; p1 < 2 * nb; ++p1) {
      int k = p1 / 2, dir = p1 % 2;  // dir 0: outgoing, 1: incoming
      size_t rbase = off[4 * k + 2 * dir], cbase = off[4 * k + 2 * dir + 1];
      int N1 = (int)(off[4 * k + 2 * dir + 1] - rbase), N2 = (int)(off[4 * k + 2 * dir + 2] - cbase);
      int np = N1 < N2 ? N1 : N2, ns = N1 < N2 ? N2 : N1;
      size_t rm = np > 0 ? np : 1;
      size_t need = rm * (2 * (size_t)np + 2 * (size_t)ns + 2) * 16 + (ns + np) * 18 + rm * 8 + 4096;
      if (p1 > p0 && bytes + need > budget) break;
      bytes += need;
      AcaProblem a;
      a.rows = didx + rbase;
      a.cols = didx + cbase;
      a.np = np;
      a.ns = ns;
      a.prim_rows = N1 < N2 ? 1 : 0;
      a.rmax = np;
      a.U = acaws.get<double2>(rm * (np ? np : 1));
      a.V = acaws.get<double2>(rm * (ns ? ns : 1));
      a.Craw = acaws.get<double2>(rm * (np ? np : 1));
      a.Lraw = acaws.get<double2>(rm * (ns ? ns : 1));
      a.line = acaws.get<double2>(ns ? ns : 1);
      a.cross = acaws.get<double2>(np ? np : 1);
      a.remP = acaws.get<char>(np ? np : 1);
      a.remS = acaws.get<char>(ns ? ns : 1);
      // rank and pivot lists go to one array for the whole batch (below)
      roff.push_back(rtot);
      rtot += 1 + 2 * rm;
      probs.push_back(a);
      res[p1].np = np;
      res[p1].ns = ns;
      res[p1].prim_rows = a.prim_rows;
    }
    int* dres = acaws.get<int>(rtot);
    for (size_t q = 0; q < probs.size(); ++q) {
      const size_t rm = probs[q].np > 0 ? probs[q].np : 1;
      probs[q].rank_out = dres + roff[q];
      probs[q].pb = dres + roff[q] + 1;
      probs[q].sb = dres + roff[q] + 1 + rm;
    }
    AcaProblem* dp = upload(scratch, probs, stream);
    launch_aca(ctx, (int)probs.size(), dp, tol, tol_robust, stream);
    // every problem's rank and pivots in one copy
    std::vector<int> hres(rtot);
    d2h(hres.data(), dres, rtot * sizeof(int), stream);
    HIP_OK(hipStreamSynchronize(stream));
    CopyBatch<double2> keepcp;
    for (int i = p0; i < p1; ++i) {
      const AcaProblem& a = probs[i - p0];
      const size_t q = (size_t)(i - p0), rm = a.np > 0 ? a.np : 1;
      const int rk = hres[roff[q]];
      if (rk < 0 || (size_t)rk > rm) throw std::runtime_error("ACA returned an impossible rank");
      res[i].rank = rk;
      res[i].pb.assign(hres.begin() + roff[q] + 1, hres.begin() + roff[q] + 1 + rk);
      res[i].sb.assign(hres.begin() + roff[q] + 1 + rm, hres.begin() + roff[q] + 1 + rm + rk);
      // keep only what the transfer operators need: Ac (incoming, N1 x rk)
      // or Ar^T (outgoing, N2 x rk), stored column-major.
      int dir = i % 2;
      bool want_cross = dir == 1 ? a.prim_rows : !a.prim_rows;
      int len = want_cross ? a.np : a.ns;
      Blk<double2> src{want_cross ? a.Craw : a.Lraw, len, rk, len > 0 ? len : 1};
      Blk<double2> dst = new_blk<double2>(acakeep, len, rk);
      keepcp.add(src, dst);
      res[i].keep = dst;
    }
    keepcp.run(scratch, stream);
    HIP_OK(hipStreamSynchronize(stream));
    acaws.reset();
    p0 = p1;
  }
  // skeleton index sets and truncation r = min(r_in, r_out)
  for (int k = 0; k < nb; ++k) {
    SkelBox& B = S[k];
    const Res& o = res[2 * k];
    const Res& in = res[2 * k + 1];
    const auto& outR = L4[4 * k + 0];
    const auto& outC = L4[4 * k + 1];
    const auto& inR = L4[4 * k + 2];
    const auto& inC = L4[4 * k + 3];
    B.out_chk.clear();
    B.out_chg.clear();
    B.in_chk.clear();
    B.in_chg.clear();
    for (int q = 0; q < o.rank; ++q) {
      int rp = o.prim_rows ? o.pb[q] : o.sb[q];
      int cp = o.prim_rows ? o.sb[q] : o.pb[q];
      B.out_chk.push_back(outR[rp]);
      B.out_chg.push_back(outC[cp]);
    }
    for (int q = 0; q < in.rank; ++q) {
      int rp = in.prim_rows ? in.pb[q] : in.sb[q];
      int cp = in.prim_rows ? in.sb[q] : in.pb[q];
      B.in_chk.push_back(inR[rp]);
      B.in_chg.push_back(inC[cp]);
    }
    int r = (int)std::min(B.in_chk.size(), B.out_chk.size());
    B.in_chk.resize(r);
    B.in_chg.resize(r);
    B.out_chk.resize(r);
    B.out_chg.resize(r);
    B.r = r;
    B.m = (int)inR.size();
  }
  // D1 = K(in_chk, in_chg), D2 = K(out_chk, out_chg)
  std::vector<int64_t> sidx;
  std::vector<size_t> soff(4 * nb + 1);
  for (int k = 0; k < nb; ++k) {
    soff[4 * k + 0] = sidx.size();
    sidx.insert(sidx.end(), S[k].in_chk.begin(), S[k].in_chk.end());
    soff[4 * k + 1] = sidx.size();
    sidx.insert(sidx.end(), S[k].in_chg.begin(), S[k].in_chg.end());
    soff[4 * k + 2] = sidx.size();
    sidx.insert(sidx.end(), S[k].out_chk.begin(), S[k].out_chk.end());
    soff[4 * k + 3] = sidx.size();
    sidx.insert(sidx.end(), S[k].out_chg.begin(), S[k].out_chg.end());
  }
  int64_t* dsidx = upload(scratch, sidx, stream);
  std::vector<BlockDesc> bd;
  std::vector<Blk<double2>> D1(nb), D2(nb), D1H(nb), AcH(nb), Br(nb);
  std::vector<int*> piv1(nb), piv2(nb);
  CopyBatch<double2> cp;
  for (int k = 0; k < nb; ++k) {
    int r = S[k].r;
    if (r == 0) continue;
    D1[k] = new_blk<double2>(scratch, r, r);
    D2[k] = new_blk<double2>(scratch, r, r);
    bd.push_back(BlockDesc{dsidx + soff[4 * k + 0], dsidx + soff[4 * k + 1], D1[k].p, r, r, r});
    bd.push_back(BlockDesc{dsidx + soff[4 * k + 2], dsidx + soff[4 * k + 3], D2[k].p, r, r, r});
  }
  BlockDesc* dbd = upload(scratch, bd, stream);
  launch_assemble(ctx, (int)bd.size(), dbd, stream);
  LUBatch<double2> lu;
  SolveBatch<double2> sv;
  for (int k = 0; k < nb; ++k) {
    int r = S[k].r;
    if (r == 0) continue;
    const Res& in = res[2 * k + 1];
    const Res& o = res[2 * k];
    int N1in = (int)L4[4 * k + 2].size();
    int N2out = (int)L4[4 * k + 1].size();
    // Ac (N1 x r_aca), column-major, ld N1: crosses if primary = rows, else lines
    Blk<double2> Ac = in.keep.sub(0, 0, N1in, r);
    // Ar^T (N2 x r_aca), column-major ld N2
    Blk<double2> ArT = o.keep.sub(0, 0, N2out, r);
    D1H[k] = new_blk<double2>(scratch, r, r);
    AcH[k] = new_blk<double2>(scratch, r, N1in);
    Br[k] = new_blk<double2>(scratch, r, N2out);
    cp.add(D1[k], D1H[k], 1);
    cp.add(Ac, AcH[k], 1);
    cp.add(ArT, Br[k], 2);
    piv1[k] = scratch.get<int>(r);
    piv2[k] = scratch.get<int>(r);
    lu.add(D1H[k], piv1[k]);
    lu.add(D2[k], piv2[k]);
  }
  cp.run(scratch, stream);
  lu.run(scratch, stream);
  for (int k = 0; k < nb; ++k) {
    if (S[k].r == 0) continue;
    sv.add(D1H[k], piv1[k], AcH[k]);
    sv.add(D2[k], piv2[k], Br[k]);
  }
  sv.run(scratch, stream);
  for (int k = 0; k < nb; ++k) {
    SkelBox& B = S[k];
    int N2out = (int)L4[4 * k + 1].size();
    B.L2P = new_blk<double2>(op, B.m, B.r);
    B.P2M = new_blk<double2>(op, B.r, N2out);
    if (B.r == 0) continue;
    cp.add(AcH[k], B.L2P, 1);
    cp.add(Br[k], B.P2M, 0);
  }
  cp.run(scratch, stream);
  HIP_OK(hipStreamSynchronize(stream));
  scratch.reset();
  acakeep.reset();
}

void Problem::assemble() {
  if (assembled) return;
  need_uniform("assembly");
  double t0 = now_s();
  sk.assign(L + 1, {});
  for (int j = 2; j <= L; ++j) sk[j].resize(1 << (2 * j));
  for (int j = L; j >= 2; --j) nca_level(j);
  HIP_OK(hipStreamSynchronize(stream));
  double t1 = now_s();
  // M2L[k][b] = K(in_chk_k, out_chg_b), b in IL(k) (FMM2DTree_find_error.hpp:1338-1361)
  for (int j = 2; j <= L; ++j) {
    int nb = 1 << (2 * j);
    std::vector<int64_t> sidx;
    std::vector<size_t> oin(nb), oout(nb);
    for (int k = 0; k < nb; ++k) {
      oin[k] = sidx.size();
      sidx.insert(sidx.end(), sk[j][k].in_chk.begin(), sk[j][k].in_chk.end());
      oout[k] = sidx.size();
      sidx.insert(sidx.end(), sk[j][k].out_chg.begin(), sk[j][k].out_chg.end());
    }
    int64_t* d = upload(scratch, sidx, stream);
    std::vector<BlockDesc> bd;
    for (int k = 0; k < nb; ++k) {
      const Lists& B = lists[j][k];
      auto addb = [&](int b) {
        Blk<double2> blk = new_blk<double2>(op, sk[j][k].r, sk[j][b].r);
        sk[j][k].M2L[b] = blk;
        if (blk.empty()) return;
        bd.push_back(BlockDesc{d + oin[k], d + oout[b], blk.p, blk.r, blk.c, blk.ld});
      };
      for (int s = 0; s < 16; ++s)
        if (B.inner[s] != -1) addb(B.inner[s]);
      for (int s = 0; s < 24; ++s)
        if (B.outer[s] != -1) addb(B.outer[s]);
    }
    BlockDesc* dbd = upload(scratch, bd, stream);
    launch_assemble(ctx, (int)bd.size(), dbd, stream);
    HIP_OK(hipStreamSynchronize(stream));
    scratch.reset();
  }
  // leaf P2P (FMM2DTree_find_error.hpp:1729-1741)
  P2P.assign(nleaf, {});
  {
    std::vector<BlockDesc> bd;
    for (int k = 0; k < nleaf; ++k) {
      const Lists& B = lists[L][k];
      auto addb = [&](int nn) {
        Blk<double2> blk = new_blk<double2>(op, R, R);
        P2P[k][nn] = blk;
        bd.push_back(BlockDesc{d_iota + (int64_t)k * R, d_iota + (int64_t)nn * R, blk.p, R, R, R});
      };
      for (int s = 0; s < 8; ++s)
        if (B.nb[s] != -1) addb(B.nb[s]);
      addb(k);
    }
    BlockDesc* dbd = upload(scratch, bd, stream);
    launch_assemble(ctx, (int)bd.size(), dbd, stream);
    HIP_OK(hipStreamSynchronize(stream));
    scratch.reset();
  }
  assembled = true;
  // the ACA workspaces are only needed while assembling
  acaws.release();
  acakeep.release();
  t_nca = t1 - t0;
  t_assemble = now_s() - t0;
}

size_t Problem::device_bytes() const {
  return persist.reserved() + op.reserved() + scratch.reserved() + acaws.reserved() + acakeep.reserved();
}

}  // namespace aifmm
