// Matrix-entry evaluation A(i,j) of the 2D Lippmann–Schwinger operator on
// gfx950, and the kernels built on it: entry lists, block assembly from
// index lists, and the batched ACA ("NCA") used to pick skeletons.
//
// A(i,j) follows LS_Assembly.hpp:474-541.  Near (source leaf adjacent to or
// equal to the target leaf): table T[slot](i%R, j%R) * kappa * kappa * q(x_i)
// (+1 on the diagonal).  Far: (i/4) * kappa^2 q(x_i) * sum_{n=-60}^{60}
// M(n+60, j%R) H_n^(1)(kappa rho) e^{i n phi}, (rho, phi) = polar(x_i - c_j).
// Entries below 1e-16 in modulus are zeroed.
#include "common.h"
#include "kernels.h"
#include "prof.h"

namespace aifmm {

__device__ __forceinline__ double hypot2(double2 a) { return hypot(a.x, a.y); }

__constant__ int c_nb_off[8][2] = {{-1, -1}, {0, -1}, {1, -1}, {1, 0}, {1, 1}, {0, 1}, {-1, 1}, {-1, 0}};

// The Graf series of a far entry, sum_{n=-60}^{60} M(n+60, col) H_n(x)
// e^{i n phi} with H_{-n} = (-1)^n H_n, in registers (no per-lane array: the
// entry, assembly, ACA and dense-product kernels all inline it).
// sum_n H_n c_n with c_n = m(60+n) w^n +- m(60-n) w^-n, w = e^{i phi},
// splits into sum_n J_n c_n and i sum_n Y_n c_n:
//  * one backward pass of Miller's recurrence yields the unnormalised J_n,
//    from which it accumulates sum_n J_n c_n (n <= 60), the normalisation
//    J_0 + 2 sum_k J_2k, and the two Neumann sums that give Y_0 and Y_1
//    (Abramowitz & Stegun 9.1.88, and its term-by-term derivative with
//    Y_1 = -Y_0'):
//      Y_0 = (2/pi) (ln(x/2) + gamma) J_0 - (4/pi) sum_k (-1)^k J_2k / k
//      Y_1 = -(2/(pi x)) J_0 + (2/pi) (ln(x/2) + gamma) J_1
//            + (2/pi) sum_k (-1)^k (J_2k-1 - J_2k+1) / k
//  * one forward pass runs Y's recurrence and accumulates sum_n Y_n c_n.
// The recurrences multiply by 1/x instead of dividing by x, and w is the
// direction (dx, dy) / rho itself instead of sincos(atan2(dy, dx)); each
// changes a term by rounding only (the series agree with scipy's Y_0, Y_1 to
// 3e-14 relative for x up to 170).
struct InvTable {
  double v[1024];
  constexpr InvTable() : v() {
    for (int i = 1; i < 1024; ++i) v[i] = 1.0 / i;
  }
};
__constant__ InvTable c_inv = InvTable();

__device__ __forceinline__ double inv_int(int m) { return m < 1024 ? c_inv.v[m] : 1.0 / m; }

__device__ __forceinline__ double2 graf_c(const double2* __restrict__ Mcol, int n, double2 pw) {
  double2 c = cmul(Mcol[SERIES + n], pw);
  if (n > 0) {
    const double2 t = cmul(Mcol[SERIES - n], cconj(pw));
    c = (n & 1) ? csub(c, t) : cadd(c, t);
  }
  return c;
}

__device__ __forceinline__ double2 graf_sum_dir(const double2* __restrict__ Mcol, double x, double2 e1) {
  constexpr double kPi = 3.14159265358979323846, kGamma = 0.57721566490153286061;
  const double rx = 1.0 / x;
  // w^60 by squaring: w^4, w^8, w^16, w^32 and their product
  double2 w60;
  {
    const double2 w2 = cmul(e1, e1), w4 = cmul(w2, w2), w8 = cmul(w4, w4), w16 = cmul(w8, w8),
                  w32 = cmul(w16, w16);
    w60 = cmul(cmul(w32, w16), cmul(w8, w4));
  }
  const double2 e1c = cconj(e1);
  int big = x > (double)SERIES ? (int)x : SERIES;
  int Mst = big + 30 + (int)sqrt(40.0 * big);
  if (Mst & 1) ++Mst;
  // jc holds the unnormalised J_n of the step's n = k - 1, jp1 J_{n+1}
  double jp1 = 0.0, jc = 1e-300;
  double sum = 0.0, s0 = 0.0, s1 = 0.0;  // J_0 + 2 sum J_2k;  Neumann sums of Y_0, Y_1
  for (int k = Mst; k > SERIES + 1; --k) {
    const double jm1 = (2.0 * k * rx) * jc - jp1;
    jp1 = jc;
    jc = jm1;
    if (fabs(jc) > 0x1p800) {
      jc *= 0x1p-800;
      jp1 *= 0x1p-800;
      sum *= 0x1p-800;
      s0 *= 0x1p-800;
      s1 *= 0x1p-800;
    }
    const int n = k - 1;
    if ((n & 1) == 0) {
      const int m = n >> 1;
      sum += 2.0 * jc;
      s0 += ((m & 1) ? -inv_int(m) : inv_int(m)) * jc;
    } else {
      const int a = (n + 1) >> 1;
      const double w = inv_int(a) + inv_int(a - 1);
      s1 += ((a & 1) ? -w : w) * jc;
    }
  }
  {
    const int e = ilogb(jc);
    const double f = ldexp(1.0, -500 - e);
    jc *= f;
    jp1 *= f;
    sum *= f;
    s0 *= f;
    s1 *= f;
  }
  double2 pw = w60;
  double accj_re = 0.0, accj_im = 0.0;  // sum J~_n c_n (unnormalised)
#pragma unroll 2
  for (int k = SERIES + 1; k >= 1; --k) {
    const double jm1 = (2.0 * k * rx) * jc - jp1;
    jp1 = jc;
    jc = jm1;
    const int n = k - 1;
    if ((n & 1) == 0) {
      if (n > 0) {
        const int m = n >> 1;
        sum += 2.0 * jc;
        s0 += ((m & 1) ? -inv_int(m) : inv_int(m)) * jc;
      }
    } else {
      const int a = (n + 1) >> 1;
      const double w = a > 1 ? inv_int(a) + inv_int(a - 1) : 1.0;
      s1 += ((a & 1) ? -w : w) * jc;
    }
    const double2 c = graf_c(Mcol, n, pw);
    accj_re += jc * c.x;
    accj_im += jc * c.y;
    if (n > 0) pw = cmul(pw, e1c);
  }
  sum += jc;
  const double rs = 1.0 / sum;
  const double j0 = jc * rs, j1 = jp1 * rs;
  const double lg = log(0.5 * x) + kGamma;
  const double y0v = (2.0 / kPi) * lg * j0 - (4.0 / kPi) * (s0 * rs);
  const double y1v = -(2.0 / kPi) * rx * j0 + (2.0 / kPi) * lg * j1 + (2.0 / kPi) * (s1 * rs);
  // forward: Y_n and w^n
  pw = mk(1.0, 0.0);
  double accy_re = 0.0, accy_im = 0.0;  // sum Y_n c_n
  double ym1 = 0.0, yc = y0v;
#pragma unroll 2
  for (int n = 0; n <= SERIES; ++n) {
    if (n == 1) {
      ym1 = y0v;
      yc = y1v;
    } else if (n > 1) {
      const double yn = (2.0 * (n - 1) * rx) * yc - ym1;
      ym1 = yc;
      yc = yn;
    }
    const double2 c = graf_c(Mcol, n, pw);
    accy_re += yc * c.x;
    accy_im += yc * c.y;
    pw = cmul(pw, e1);
  }
  // sum H_n c_n = sum J_n c_n + i sum Y_n c_n
  return mk(accj_re * rs - accy_im, accj_im * rs + accy_re);
}

// The near table of target leaf li for source leaf lj (adaptive), or null
// when the pair is far.
__device__ __forceinline__ const double2* adaptive_near_table(const EntryCtx& C, int li, int lj) {
  for (int e = C.near_off[li]; e < C.near_off[li + 1]; ++e)
    if (C.near_leaf[e] == lj) return C.Ttab + (size_t)C.near_tab[e] * C.R * C.R;
  return nullptr;
}

// getMatrixEntry between leaves of an adaptive tree (LS_Assembly.hpp:474-541):
// the near table of the pair's neighbour type at the target's level, or the
// Graf series with the far table of the source leaf's level about its centre.
__device__ double2 ls_entry_adaptive(const EntryCtx& C, int64_t i, int64_t j) {
#pragma clang fp contract(off)
  const int R = C.R;
  const int li = (int)(i / R), ii = (int)(i - (int64_t)li * R);
  const int lj = (int)(j / R), jj = (int)(j - (int64_t)lj * R);
  const double qi = C.q[i];
  const double2* T = adaptive_near_table(C, li, lj);
  if (T) {
    double2 e = T[ii + (size_t)R * jj];
    if (C.find_phi) {
      e = cscale(e, C.kappa);
      e = cscale(e, C.kappa);
      e = cscale(e, qi);
      if (i == j) e.x = 1.0 + e.x;
    }
    if (hypot(e.x, e.y) < 1e-16) e = mk(0.0, 0.0);
    return e;
  }
  const double ddy = C.py[i] - C.lcy[lj], ddx = C.px[i] - C.lcx[lj];
  const double rho = sqrt(ddy * ddy + ddx * ddx), ir = 1.0 / rho;
  double2 s = graf_sum_dir(C.Mlev + (size_t)C.leaf_lv[lj] * NTERM * R + (size_t)NTERM * jj, C.kappa * rho,
                           mk(ddx * ir, ddy * ir));
  s = cmul(s, mk(0.0, 0.25));
  if (C.find_phi) {
    s = cscale(s, C.kappa * C.kappa * qi);
    if (i == j) s.x += 1.0;
  }
  if (hypot(s.x, s.y) < 1e-16) s = mk(0.0, 0.0);
  return s;
}

__device__ double2 ls_entry(const EntryCtx& C, int64_t i, int64_t j) {
#pragma clang fp contract(off)
  if (C.adaptive) return ls_entry_adaptive(C, i, j);
  int R = C.R;
  int li = (int)(i / R), ii = (int)(i - (int64_t)li * R);
  int lj = (int)(j / R), jj = (int)(j - (int64_t)lj * R);
  int xi, yi, xj, yj;
  box_xy(C.L, li, xi, yi);
  box_xy(C.L, lj, xj, yj);
  int dx = xj - xi, dy = yj - yi;
  double qi = C.q[i];
  if (dx >= -1 && dx <= 1 && dy >= -1 && dy <= 1) {
    int nb = 8;
    if (dx || dy)
      for (int s = 0; s < 8; ++s)
        if (c_nb_off[s][0] == dx && c_nb_off[s][1] == dy) nb = s;
    double2 e = C.T[(size_t)nb * R * R + ii + (size_t)R * jj];
    if (C.find_phi) {
      e = cscale(e, C.kappa);
      e = cscale(e, C.kappa);
      e = cscale(e, qi);
      if (i == j) e.x = 1.0 + e.x;
    }
    if (hypot(e.x, e.y) < 1e-16) e = mk(0.0, 0.0);
    return e;
  }
  double cx = -C.Ldom + (2 * xj + 1) * C.rleaf, cy = -C.Ldom + (2 * yj + 1) * C.rleaf;
  double ddy = C.py[i] - cy, ddx = C.px[i] - cx;
  const double rho = sqrt(ddy * ddy + ddx * ddx), ir = 1.0 / rho;
  double2 s = graf_sum_dir(C.M + (size_t)NTERM * jj, C.kappa * rho, mk(ddx * ir, ddy * ir));
  s = cmul(s, mk(0.0, 0.25));
  if (C.find_phi) {
    s = cscale(s, C.kappa * C.kappa * qi);
    if (i == j) s.x += 1.0;
  }
  if (hypot(s.x, s.y) < 1e-16) s = mk(0.0, 0.0);
  return s;
}

// lane l's value of v in every lane (l uniform)
__device__ __forceinline__ double readlane1(double v, int l) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readlane((int)b, l), hi = __builtin_amdgcn_readlane((int)(b >> 32), l);
  return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}
__device__ __forceinline__ double2 readlane2(double2 v, int l) { return mk(readlane1(v.x, l), readlane1(v.y, l)); }

__device__ __forceinline__ double2 near_fma(double2 a, double2 b, double2 c) {
  c.x = fma(a.x, b.x, c.x);
  c.x = fma(-a.y, b.y, c.x);
  c.y = fma(a.x, b.y, c.y);
  c.y = fma(a.y, b.x, c.y);
  return c;
}

// NEAR_BOXES leaf boxes per workgroup (grid row = problem); wave w takes
// neighbour slots w, w + 4, w + 8; lane = row of a box (R <= 64).  Neighbour
// slot s always uses table s (the slots follow the tables' offset order,
// checked on the host), so a wave loads each table entry once and applies it
// to all its boxes.  The near entries are kappa^2 q_i T_s[ii, jj] (+ 1 on the
// diagonal): with D.T = kappa^2 T, each row sums T_s x_nn over its slots and
// is scaled by q_i once, plus x_i.  (The assembled blocks round each entry
// and zero those below 1e-16; the product differs from theirs at rounding
// level only.)  Measured on MI355X at N = 147,456 (refined solve): 1 box per
// workgroup 6.80 ms, 2 boxes 6.71 ms, 4 boxes 6.79 ms, 8 boxes 6.92 ms.
constexpr int NEAR_BOXES = 2;

// CC right-hand-side columns per workgroup (grid z = column groups): each
// table entry loaded once is applied to every column of the group; every
// column's sums are the same operations in the same order for any CC.
template <int CC>
__global__ __launch_bounds__(256) void k_near(const NearDesc* __restrict__ ds) {
  const NearDesc D = ds[blockIdx.y];
  const int k0 = blockIdx.x * NEAR_BOXES;
  const int c0 = blockIdx.z * CC;
  if (k0 >= D.nbox || c0 >= D.nrhs) return;
  const int nc = D.nrhs - c0 < CC ? D.nrhs - c0 : CC;
  const int R = D.R, lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  __shared__ double2 red[4][NEAR_BOXES][CC][64];
  double2 acc[NEAR_BOXES][CC];
#pragma unroll
  for (int b = 0; b < NEAR_BOXES; ++b)
#pragma unroll
    for (int u = 0; u < CC; ++u) acc[b][u] = mk(0.0, 0.0);
  if (lane < R)
    for (int sl = w; sl < 9; sl += 4) {
      const double2* xb[NEAR_BOXES];
#pragma unroll
      for (int b = 0; b < NEAR_BOXES; ++b) {
        const int k = k0 + b;
        const int nn = k < D.nbox ? D.nbr[(size_t)k * 9 + sl] : -1;
        xb[b] = nn < 0 ? nullptr : D.x + (size_t)D.N * c0 + (size_t)nn * R;
      }
      // each neighbour segment in one coalesced load (lane jj holds x[jj]),
      // handed to every lane by readlane as jj advances
      double2 xv[NEAR_BOXES][CC];
#pragma unroll
      for (int b = 0; b < NEAR_BOXES; ++b)
#pragma unroll
        for (int u = 0; u < CC; ++u) xv[b][u] = xb[b] && u < nc ? xb[b][(size_t)D.N * u + lane] : mk(0.0, 0.0);
      const double2* Tt = D.T + (size_t)sl * R * R + lane;
      for (int jj = 0; jj < R; ++jj) {
        const double2 e = Tt[(size_t)R * jj];
#pragma unroll
        for (int b = 0; b < NEAR_BOXES; ++b)
          if (xb[b]) {
#pragma unroll
            for (int u = 0; u < CC; ++u)
              if (u < nc) acc[b][u] = near_fma(e, readlane2(xv[b][u], jj), acc[b][u]);
          }
      }
    }
#pragma unroll
  for (int b = 0; b < NEAR_BOXES; ++b)
#pragma unroll
    for (int u = 0; u < CC; ++u) red[w][b][u][lane] = acc[b][u];
  __syncthreads();
  // wave w finishes boxes w, w + 4 of the group
  for (int b = w; b < NEAR_BOXES; b += 4) {
    const int k = k0 + b;
    if (k < D.nbox && lane < R) {
      const int64_t i = (int64_t)k * R + lane;
      for (int u = 0; u < nc; ++u) {
        double2 t = red[0][b][u][lane];
        t = cadd(t, red[1][b][u][lane]);
        t = cadd(t, red[2][b][u][lane]);
        t = cadd(t, red[3][b][u][lane]);
        const size_t o = (size_t)D.N * (c0 + u) + i;
        double2 v = cscale(t, D.q[i]);
        v = cadd(v, D.x[o]);  // the identity
        D.y[o] = cadd(D.y[o], v);
      }
    }
  }
}

void launch_near(int nmem, int max_box, const NearDesc* d, hipStream_t s, int max_nrhs) {
  if (!nmem || max_box < 1 || max_nrhs < 1) return;
  const int gx = (max_box + NEAR_BOXES - 1) / NEAR_BOXES;
  // measured at N = 589,824 with 256 columns: 1 column per workgroup 37.0 ms,
  // 4 columns 32.2 ms, 8 columns 42.6 ms (fewer waves per SIMD)
  if (max_nrhs >= 4)
    hipLaunchKernelGGL(k_near<4>, dim3(gx, nmem, (max_nrhs + 3) / 4), dim3(256), 0, s, d);
  else
    hipLaunchKernelGGL(k_near<1>, dim3(gx, nmem, max_nrhs), dim3(256), 0, s, d);
  HIP_OK(hipGetLastError());
}

__global__ void k_entries(EntryCtx C, int64_t n, const int64_t* __restrict__ I, const int64_t* __restrict__ Jv,
                          double2* __restrict__ out) {
  int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (t < n) out[t] = ls_entry(C, I[t], Jv[t]);
}

// One workgroup per block descriptor: out(r, c) = A(rows[r], cols[c]),
// column-major with leading dimension ld.
__global__ void k_assemble(EntryCtx C, const BlockDesc* __restrict__ d) {
  BlockDesc b = d[blockIdx.x];
  int tot = b.nr * b.nc;
  for (int t = threadIdx.x; t < tot; t += blockDim.x) {
    int r = t % b.nr, c = t / b.nr;
    b.out[r + (size_t)b.ld * c] = ls_entry(C, b.rows[r], b.cols[c]);
  }
}

// ---------------------------------------------------------------- ACA
// LowRank::ACA_only_nodes (ACA.hpp:368-577) in "primary/secondary" form:
// the primary index set P is the shorter dimension (rows when N1 < N2, else
// columns); a "line" is A restricted to one P index (length ns), a "cross"
// is A restricted to one S index (length np).  One workgroup per problem.
constexpr int ACA_THREADS = 256;

struct RedPair {
  double v;
  int i;
};

__device__ __forceinline__ RedPair red_max(RedPair a, RedPair b) {
  // larger magnitude wins; ties go to the lower index (first strict max)
  if (b.i < 0) return a;
  if (a.i < 0) return b;
  if (b.v > a.v || (b.v == a.v && b.i < a.i)) return b;
  return a;
}

__device__ RedPair block_argmax(RedPair p, RedPair* sh) {
  int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  for (int o = 32; o > 0; o >>= 1) {
    RedPair q;
    q.v = __shfl_down(p.v, o);
    q.i = __shfl_down(p.i, o);
    p = red_max(p, q);
  }
  if (lane == 0) sh[w] = p;
  __syncthreads();
  if (threadIdx.x == 0) {
    RedPair r = sh[0];
    for (int k = 1; k < ACA_THREADS / 64; ++k) r = red_max(r, sh[k]);
    sh[4] = r;
  }
  __syncthreads();
  RedPair r = sh[4];
  __syncthreads();
  return r;
}

__device__ double block_sum(double v, double* sh) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o);
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
  __syncthreads();
  double r = 0;
  for (int k = 0; k < ACA_THREADS / 64; ++k) r += sh[k];
  __syncthreads();
  return r;
}

// First allowed index whose modulus is within a relative 1e-12 of the
// maximum modulus (maxAbsVector, ACA.hpp:27-40, with ties taken robustly:
// mirror-symmetric geometry produces exact ties that rounding would break
// arbitrarily).  -1 when nothing is allowed.
constexpr double ACA_TIE = 1e-12;
__device__ RedPair tol_argmax(const double2* v, const char* allowed, int n, RedPair* sh, double* shd) {
  double mx = -1.0;
  for (int t = threadIdx.x; t < n; t += blockDim.x)
    if (!allowed || allowed[t]) mx = fmax(mx, hypot(v[t].x, v[t].y));
  for (int o = 32; o > 0; o >>= 1) mx = fmax(mx, __shfl_xor(mx, o));
  if ((threadIdx.x & 63) == 0) shd[threadIdx.x >> 6] = mx;
  __syncthreads();
  mx = shd[0];
  for (int k = 1; k < ACA_THREADS / 64; ++k) mx = fmax(mx, shd[k]);
  __syncthreads();
  if (mx < 0) return RedPair{-1.0, -1};
  double thr = mx * (1.0 - ACA_TIE);
  int best = 0x7fffffff;
  for (int t = threadIdx.x; t < n; t += blockDim.x)
    if ((!allowed || allowed[t]) && hypot(v[t].x, v[t].y) >= thr) {
      best = t;
      break;
    }
  for (int o = 32; o > 0; o >>= 1) best = min(best, __shfl_xor(best, o));
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6].i = best;
  __syncthreads();
  for (int k = 0; k < ACA_THREADS / 64; ++k) best = min(best, sh[k].i);
  __syncthreads();
  return RedPair{mx, best == 0x7fffffff ? -1 : best};
}

__device__ __forceinline__ double2 aca_entry(const EntryCtx& C, const AcaProblem& p, int pi, int si) {
  // primary index pi, secondary index si -> A(row, col)
  if (p.prim_rows) return ls_entry(C, p.rows[pi], p.cols[si]);
  return ls_entry(C, p.rows[si], p.cols[pi]);
}

__global__ __launch_bounds__(ACA_THREADS) void k_aca(EntryCtx C, const AcaProblem* __restrict__ probs, double tol,
                                                     double tol_robust) {
#pragma clang fp contract(off)
  const AcaProblem p = probs[blockIdx.x];
  const int np = p.np, ns = p.ns;
  __shared__ RedPair shp[8];
  __shared__ double shd[8];
  __shared__ int s_flag, s_pi, s_si, s_rank;
  __shared__ double s_normS, s_prev;
  __shared__ double2 s_piv;
  double2* line = p.line;    // ns
  double2* crossr = p.cross; // np (residual)
  char* remP = p.remP;
  char* remS = p.remS;
  for (int t = threadIdx.x; t < np; t += blockDim.x) remP[t] = 1;
  for (int t = threadIdx.x; t < ns; t += blockDim.x) remS[t] = 1;
  if (threadIdx.x == 0) s_rank = 0;
  __syncthreads();
  if (np == 0 || ns == 0) {
    if (threadIdx.x == 0) *p.rank_out = 0;
    return;
  }
  // first primary index whose line has a pivot above tol_robust
  int pi = -1, si = 0;
  for (int l = 0; l < np; ++l) {
    for (int t = threadIdx.x; t < ns; t += blockDim.x) line[t] = aca_entry(C, p, l, t);
    __syncthreads();
    RedPair r = tol_argmax(line, remS, ns, shp, shd);
    if (r.v > tol_robust) {
      pi = l;
      si = r.i;
      break;
    }
  }
  if (pi < 0) {
    if (threadIdx.x == 0) *p.rank_out = 0;
    return;
  }
  const int mn = np < ns ? np : ns;
  double2 piv = line[si];
  // v_0 = line / pivot ; cross_0 = raw cross at si ; u_0 = cross_0
  for (int t = threadIdx.x; t < ns; t += blockDim.x) {
    double2 a = line[t];
    double d = cabs2(piv);
    p.V[t] = mk((a.x * piv.x + a.y * piv.y) / d, (a.y * piv.x - a.x * piv.y) / d);
    p.Lraw[t] = a;
  }
  for (int t = threadIdx.x; t < np; t += blockDim.x) {
    double2 e = aca_entry(C, p, t, si);
    p.Craw[t] = e;
    p.U[t] = e;
    crossr[t] = e;
  }
  if (threadIdx.x == 0) {
    p.pb[0] = pi;
    p.sb[0] = si;
    remS[si] = 0;
    remP[pi] = 0;
    s_normS = 0.0;
    s_prev = 1.7976931348623157e308;
    s_rank = 1;
  }
  __syncthreads();
  int rank = 1;
  // next primary candidate from the cross
  {
    RedPair r = tol_argmax(crossr, remP, np, shp, shd);
    pi = r.i < 0 ? 0 : r.i;
  }
  double lastpiv = hypot2(piv);
  while (true) {
    double cpi = hypot2(crossr[pi]);
    if (!(lastpiv > tol && cpi > tol && fabs(s_prev - s_normS) >= tol * s_prev && rank < mn)) break;
    if (rank >= p.rmax) break;
    // residual line at pi
    for (int t = threadIdx.x; t < ns; t += blockDim.x) {
      double2 e = aca_entry(C, p, pi, t);
      p.Lraw[(size_t)rank * ns + t] = e;
      for (int l = 0; l < rank; ++l) e = csub(e, cmul(p.U[(size_t)l * np + pi], p.V[(size_t)l * ns + t]));
      line[t] = e;
    }
    __syncthreads();
    RedPair r = tol_argmax(line, remS, ns, shp, shd);
    si = r.i < 0 ? 0 : r.i;
    piv = line[si];
    double2* Vn = p.V + (size_t)rank * ns;
    double vv = 0;
    for (int t = threadIdx.x; t < ns; t += blockDim.x) {
      double2 a = line[t];
      double d = cabs2(piv);
      double2 v = mk((a.x * piv.x + a.y * piv.y) / d, (a.y * piv.x - a.x * piv.y) / d);
      Vn[t] = v;
      vv += cabs2(v);
    }
    double2* Un = p.U + (size_t)rank * np;
    double uu = 0;
    for (int t = threadIdx.x; t < np; t += blockDim.x) {
      double2 e = aca_entry(C, p, t, si);
      p.Craw[(size_t)rank * np + t] = e;
      for (int l = 0; l < rank; ++l) e = csub(e, cmul(p.V[(size_t)l * ns + si], p.U[(size_t)l * np + t]));
      Un[t] = e;
      crossr[t] = e;
      uu += cabs2(e);
    }
    __syncthreads();
    double unorm = sqrt(block_sum(uu, shd));
    double vnorm = sqrt(block_sum(vv, shd));
    if (unorm < tol_robust || vnorm < tol_robust) break;
    // normS update: normS^2 + (|u||v|)^2 + sum_l 2|(U_l^H u)(V_l^H v)| over
    // l = 0..rank (the new pair included, as the reference does).  One wave
    // per l, no workgroup barriers inside the loop.
    double accw = 0;
    {
      int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
      for (int l = w; l <= rank; l += ACA_THREADS / 64) {
        double2 du = mk(0, 0), dv = mk(0, 0);
        for (int t = lane; t < np; t += 64) du = cadd(du, cmul(cconj(p.U[(size_t)l * np + t]), Un[t]));
        for (int t = lane; t < ns; t += 64) dv = cadd(dv, cmul(cconj(p.V[(size_t)l * ns + t]), Vn[t]));
        for (int o = 32; o > 0; o >>= 1) {
          du.x += __shfl_xor(du.x, o);
          du.y += __shfl_xor(du.y, o);
          dv.x += __shfl_xor(dv.x, o);
          dv.y += __shfl_xor(dv.y, o);
        }
        double2 pr = cmul(du, dv);
        accw += 2.0 * hypot(pr.x, pr.y);
      }
      if (lane == 0) shd[w] = accw;
      __syncthreads();
    }
    double accn = 0;
    for (int k = 0; k < ACA_THREADS / 64; ++k) accn += shd[k];
    __syncthreads();
    if (threadIdx.x == 0) {
      p.pb[rank] = pi;
      p.sb[rank] = si;
      remS[si] = 0;
      remP[pi] = 0;
    }
    ++rank;
    if (rank != 2) s_prev = s_normS;  // uniform across threads
    double ns2 = s_normS * s_normS + (unorm * vnorm) * (unorm * vnorm) + accn;
    __syncthreads();
    if (threadIdx.x == 0) s_normS = sqrt(ns2);
    lastpiv = hypot2(piv);
    __syncthreads();
    RedPair r2 = tol_argmax(crossr, remP, np, shp, shd);
    pi = r2.i < 0 ? 0 : r2.i;
  }
  if (threadIdx.x == 0) *p.rank_out = rank;
}

// ------------------------------------------------------------ launchers
// ---------------------------------------------------------------- dense product
// y = A x with every entry of A as ls_entry, over all leaf pairs: the
// generator's dense rhs = A x (FMM2DTree_gen_rhs_x.hpp:2808-2903).  A far
// entry from source leaf l is (i/4) kappa^2 q_i sum_n M(n, jj) H_n e^{i n phi}
// with (rho, phi) taken from the leaf's centre, so the leaf's contribution to
// row i is (i/4) kappa^2 q_i sum_n m_l(n) H_n e^{i n phi}, with the leaf
// moments m_l = M x_l (121 values, the layout of a table column).  One Graf
// series per (target point, source leaf) instead of one per entry.  The near
// 3 x 3 neighbourhood is summed entry by entry, each entry formed and zeroed
// below 1e-16 as ls_entry does; far entries are summed unzeroed (an entry
// below 1e-16 changes its row sum by less than 1e-16 |x_j|).
constexpr int DMV_T = 256;

// m[l * NTERM + n] = sum_jj M(n, jj) x[l R + jj]
__global__ __launch_bounds__(128) void k_dmv_moments(EntryCtx C, const double2* __restrict__ x,
                                                     double2* __restrict__ m) {
  const int l = blockIdx.x, n = threadIdx.x;
  if (n >= NTERM) return;
  const double2* xl = x + (size_t)l * C.R;
  // the far table of the leaf's level (adaptive) or of the leaf level
  const double2* M = C.adaptive ? C.Mlev + (size_t)C.leaf_lv[l] * NTERM * C.R : C.M;
  double2 acc = mk(0.0, 0.0);
  for (int jj = 0; jj < C.R; ++jj) acc = cadd(acc, cmul(M[n + (size_t)NTERM * jj], xl[jj]));
  m[(size_t)l * NTERM + n] = acc;
}

// grid (row tiles of DMV_T target points, leaf chunks): part[c N + i] = the
// Graf sums of target i over the far leaves of chunk c
__global__ __launch_bounds__(DMV_T) void k_dmv_far(EntryCtx C, const double2* __restrict__ m,
                                                   double2* __restrict__ part, int64_t N, int nleaf, int lpc) {
  const int64_t i = (int64_t)blockIdx.x * DMV_T + threadIdx.x;
  const int c = blockIdx.y;
  if (i >= N) return;
  const int l0 = c * lpc, l1 = min(nleaf, l0 + lpc);
  const int li = (int)(i / C.R);
  int xi, yi;
  box_xy(C.L, li, xi, yi);
  const double px = C.px[i], py = C.py[i];
  // adaptive: getIfNeighbor's near test between the two leaves
  const double ai = C.adaptive ? 3.0 * C.rad[C.leaf_lv[li]] - 1e-16 : 0.0;
  const double cxi = C.adaptive ? C.lcx[li] : 0.0, cyi = C.adaptive ? C.lcy[li] : 0.0;
  double2 acc = mk(0.0, 0.0);
  for (int l = l0; l < l1; ++l) {
    double cx, cy;
    if (C.adaptive) {
      cx = C.lcx[l];
      cy = C.lcy[l];
      const double lim = ai + C.rad[C.leaf_lv[l]];
      if (!(fabs(cxi - cx) >= lim || fabs(cyi - cy) >= lim)) continue;
    } else {
      int xj, yj;
      box_xy(C.L, l, xj, yj);
      if (xj - xi >= -1 && xj - xi <= 1 && yj - yi >= -1 && yj - yi <= 1) continue;
      cx = -C.Ldom + (2 * xj + 1) * C.rleaf;
      cy = -C.Ldom + (2 * yj + 1) * C.rleaf;
    }
    const double ddy = py - cy, ddx = px - cx;
    const double rho = sqrt(ddy * ddy + ddx * ddx);
    const double ir = 1.0 / rho;
    acc = cadd(acc, graf_sum_dir(m + (size_t)NTERM * l, C.kappa * rho, mk(ddx * ir, ddy * ir)));
  }
  part[(size_t)c * N + i] = acc;
}

// y_i = (i/4) kappa^2 q_i sum_c part[c N + i] (chunks in order) + the near
// entries of row i times x
__global__ __launch_bounds__(DMV_T) void k_dmv_sum(EntryCtx C, const double2* __restrict__ x,
                                                   const double2* __restrict__ part, int nchunk, double2* __restrict__ y,
                                                   int64_t N) {
  const int64_t i = (int64_t)blockIdx.x * DMV_T + threadIdx.x;
  if (i >= N) return;
  const int R = C.R, li = (int)(i / R), ii = (int)(i - (int64_t)li * R);
  double2 far = mk(0.0, 0.0);
  for (int c = 0; c < nchunk; ++c) far = cadd(far, part[(size_t)c * N + i]);
  const double qi = C.q[i];
  double2 acc = cmul(far, mk(0.0, 0.25));
  if (C.find_phi) acc = cscale(acc, C.kappa * C.kappa * qi);
  if (C.adaptive) {
    for (int e = C.near_off[li]; e < C.near_off[li + 1]; ++e) {
      const int lj = C.near_leaf[e];
      const double2* Tt = C.Ttab + (size_t)C.near_tab[e] * R * R + ii;
      const double2* xl = x + (size_t)lj * R;
      for (int jj = 0; jj < R; ++jj) {
        double2 v = Tt[(size_t)R * jj];
        if (C.find_phi) {
          v = cscale(v, C.kappa);
          v = cscale(v, C.kappa);
          v = cscale(v, qi);
          if (lj == li && jj == ii) v.x = 1.0 + v.x;
        }
        if (hypot(v.x, v.y) < 1e-16) v = mk(0.0, 0.0);
        acc = cadd(acc, cmul(v, xl[jj]));
      }
    }
    y[i] = acc;
    return;
  }
  int xi, yi;
  box_xy(C.L, li, xi, yi);
  const int side = 1 << C.L;
  for (int nb = 0; nb < 9; ++nb) {
    const int dx = nb < 8 ? c_nb_off[nb][0] : 0, dy = nb < 8 ? c_nb_off[nb][1] : 0;
    if (xi + dx < 0 || xi + dx >= side || yi + dy < 0 || yi + dy >= side) continue;
    const int lj = xy_box(C.L, xi + dx, yi + dy);
    const double2* Tt = C.T + (size_t)nb * R * R + ii;
    const double2* xl = x + (size_t)lj * R;
    for (int jj = 0; jj < R; ++jj) {
      double2 e = Tt[(size_t)R * jj];
      if (C.find_phi) {
        e = cscale(e, C.kappa);
        e = cscale(e, C.kappa);
        e = cscale(e, qi);
        if (nb == 8 && jj == ii) e.x = 1.0 + e.x;
      }
      if (hypot(e.x, e.y) < 1e-16) e = mk(0.0, 0.0);
      acc = cadd(acc, cmul(e, xl[jj]));
    }
  }
  y[i] = acc;
}

int dmv_chunks(int64_t N, int nleaf) {
  const int64_t tiles = (N + DMV_T - 1) / DMV_T;
  int64_t c = (4096 + tiles - 1) / tiles;
  return (int)(c < 1 ? 1 : c > nleaf ? nleaf : c);
}

void launch_dense_matvec(const EntryCtx& C, int64_t N, const double2* x, double2* y, double2* scratch, int nchunk,
                         hipStream_t s, bool profile) {
  const int nleaf = (int)(N / C.R);
  if (nchunk < 1 || nchunk > nleaf) throw std::runtime_error("dense product: bad leaf-chunk count");
  double2* m = scratch;
  double2* part = scratch + (size_t)NTERM * nleaf;
  const int lpc = (nleaf + nchunk - 1) / nchunk;
  const unsigned tiles = (unsigned)((N + DMV_T - 1) / DMV_T);
  {
    ProfScope ps(profile ? s : nullptr, profile ? "k_dmv_moments" : nullptr,
                 16.0 * ((double)NTERM * C.R + N + (double)NTERM * nleaf), 8.0 * NTERM * (double)N);
    hipLaunchKernelGGL(k_dmv_moments, dim3(nleaf), dim3(128), 0, s, C, x, m);
    HIP_OK(hipGetLastError());
  }
  {
    // flops: an estimate per (point, far leaf) pair of the Graf series (61
    // combined terms of about 29 flops, Y recurrence included) and Miller's
    // recurrence for J (about 150 steps of 3 flops): 2,200
    ProfScope ps(profile ? s : nullptr, profile ? "k_dmv_far" : nullptr,
                 16.0 * ((double)NTERM * nleaf + (double)nchunk * N), 2200.0 * (double)N * (nleaf - 9));
    hipLaunchKernelGGL(k_dmv_far, dim3(tiles, nchunk), dim3(DMV_T), 0, s, C, m, part, N, nleaf, lpc);
    HIP_OK(hipGetLastError());
  }
  {
    ProfScope ps(profile ? s : nullptr, profile ? "k_dmv_sum" : nullptr, 16.0 * ((double)nchunk * N + 10.0 * N),
                 8.0 * 9 * C.R * (double)N);
    hipLaunchKernelGGL(k_dmv_sum, dim3(tiles), dim3(DMV_T), 0, s, C, x, part, nchunk, y, N);
    HIP_OK(hipGetLastError());
  }
}

void launch_entries(const EntryCtx& C, int64_t n, const int64_t* I, const int64_t* J, double2* out, hipStream_t s) {
  if (n == 0) return;
  int th = 128;
  hipLaunchKernelGGL(k_entries, dim3((unsigned)((n + th - 1) / th)), dim3(th), 0, s, C, n, I, J, out);
  HIP_OK(hipGetLastError());
}

void launch_assemble(const EntryCtx& C, int nblk, const BlockDesc* d, hipStream_t s) {
  if (nblk == 0) return;
  hipLaunchKernelGGL(k_assemble, dim3(nblk), dim3(256), 0, s, C, d);
  HIP_OK(hipGetLastError());
}

void launch_aca(const EntryCtx& C, int nprob, const AcaProblem* d, double tol, double tol_robust, hipStream_t s) {
  if (nprob == 0) return;
  hipLaunchKernelGGL(k_aca, dim3(nprob), dim3(ACA_THREADS), 0, s, C, d, tol, tol_robust);
  HIP_OK(hipGetLastError());
}

}  // namespace aifmm
