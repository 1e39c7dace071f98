// Precomputation tables on the device: adaptive 10x10 Gauss–Legendre
// quadrature, one wave64 per integral, with a depth-first panel stack in LDS.
//
// Rule (LS/domain2D.hpp:38-138): split a panel into 4 children (0=(-,-),
// 1=(+,-), 2=(+,+), 3=(-,+)); accept the PARENT value when
// |I_parent - sum I_children| < eps (absolute); the integral is the sum of
// accepted panels in depth-first order.
//   far  C(n,l)   = int_{[-1,1]^2} J_n(kappa r |s|) e^{-i n theta_s} T_a(s1) T_b(s2) ds
//                   (FarFieldInteraction.hpp:62-74)
//   near V(p,l)   = int_{[-1,1]^2} H_0(kappa r_B |X_p - s|) T_a(s1) T_b(s2) ds,
//                   X_p = (r_A xi_p - c_B) / r_B (NearFieldInteraction1.hpp:98-128)
#include "common.h"
#include "kernels.h"

namespace aifmm {

// J_n(z) for one order n >= 0 by Miller's backward recurrence, normalised by
// J_0 + 2 sum J_2k = 1 (power-of-two rescaling only).
__device__ double besselJn_single(int n, double z) {
#pragma clang fp contract(off)
  int big = z > (double)n ? (int)z : n;
  int Mst = big + 30 + (int)sqrt(40.0 * (big > 1 ? big : 1));
  if (Mst & 1) ++Mst;
  double jp1 = 0.0, jc = 1e-300, sum = 0.0, jn = 0.0;
  for (int k = Mst; k >= 1; --k) {
    double jm1 = (2.0 * k / z) * jc - jp1;
    jp1 = jc;
    jc = jm1;
    if (fabs(jc) > 0x1p800) {
      jc *= 0x1p-800;
      jp1 *= 0x1p-800;
      sum *= 0x1p-800;
      jn *= 0x1p-800;
    }
    if (k - 1 == n) jn = jc;
    if (k - 1 > 0 && ((k - 1) & 1) == 0) sum += 2.0 * jc;
  }
  sum += jc;
  return jn / sum;
}

__device__ __forceinline__ double chebT(double x, int n) { return cos(n * acos(x)); }

__device__ __forceinline__ double2 integrand(const QuadCtx& Q, const QuadTask& t, double x, double y) {
#pragma clang fp contract(off)
  int a = Q.basis_a[t.l], b = Q.basis_b[t.l];
  double ta = chebT(x, a), tb = chebT(y, b);
  const QuadGeom& G = Q.geom[t.g];
  if (t.kind == 0) {
    double R2 = x * x + y * y;
    double nx = sqrt(R2);
    double arg = G.kr * nx;
    double th = atan2(y, x);
    int an = t.n < 0 ? -t.n : t.n;
    double jv = besselJn_single(an, arg);
    if (t.n < 0 && (an & 1)) jv = -jv;
    double s, c;
    sincos(-(double)t.n * th, &s, &c);
    double2 v = mk(jv * c, jv * s);
    v = cscale(v, ta);
    v = cscale(v, tb);
    return v;
  } else {
    // the target node r_A xi + 0 (box A at the origin), then
    // (node - c_B) / r_B, in the reference's order of operations
    double Xx = (G.ra * Q.cheb1d[t.p % Q.n_cheb] + 0.0 - G.cx) / G.rb;
    double Xy = (G.ra * Q.cheb1d[t.p / Q.n_cheb] + 0.0 - G.cy) / G.rb;
    double R2 = (Xx - x) * (Xx - x) + (Xy - y) * (Xy - y);
    double R = sqrt(R2);
    double arg = G.kr * R;
    double2 v = mk(j0(arg), y0(arg));
    v = cscale(v, ta);
    v = cscale(v, tb);
    return v;
  }
}

constexpr int QSTK = 192;
constexpr int QMAXDEPTH = 48;

struct Panel {
  double xc, yc, h;
  double2 I;
  int depth;
};

// Integrals of np panels (np = 1 or 4) of half-width h centred at (cx[c], cy[c]).
__device__ void eval_panels(const QuadCtx& Q, const QuadTask& t, int np, const double* cx, const double* cy, double h,
                            double2* out) {
  int lane = threadIdx.x;
  double2 acc[4] = {mk(0, 0), mk(0, 0), mk(0, 0), mk(0, 0)};
  for (int pnt = lane; pnt < np * 100; pnt += 64) {
    int c = pnt / 100, mm = (pnt % 100) / 10, nn = pnt % 10;
    double x = cx[c] + h * Q.gl_t[mm], y = cy[c] + h * Q.gl_t[nn];
    double2 f = integrand(Q, t, x, y);
    double w = Q.gl_w[mm] * Q.gl_w[nn];
    acc[c] = cadd(acc[c], cscale(f, w));
  }
  for (int c = 0; c < np; ++c) {
    double2 v = acc[c];
    for (int o = 32; o > 0; o >>= 1) {
      v.x += __shfl_xor(v.x, o);
      v.y += __shfl_xor(v.y, o);
    }
    out[c] = cscale(v, h * h);
  }
}

__global__ __launch_bounds__(64) void k_quad(QuadCtx Q, const QuadTask* __restrict__ tasks, double2* __restrict__ out,
                                             int* __restrict__ depth_flag) {
  const QuadTask t = tasks[blockIdx.x];
  __shared__ Panel stk[QSTK];
  int sp = 0;
  double2 total = mk(0, 0);
  {
    double cx = 0, cy = 0;
    double2 I0;
    eval_panels(Q, t, 1, &cx, &cy, 1.0, &I0);
    if (threadIdx.x == 0) stk[0] = Panel{0.0, 0.0, 1.0, I0, 0};
    sp = 1;
  }
  __syncthreads();
  bool capped = false;
  while (sp > 0) {
    Panel P = stk[sp - 1];
    --sp;
    __syncthreads();
    double h = 0.5 * P.h;
    double cx[4] = {P.xc - h, P.xc + h, P.xc + h, P.xc - h};
    double cy[4] = {P.yc - h, P.yc - h, P.yc + h, P.yc + h};
    double2 Ic[4];
    eval_panels(Q, t, 4, cx, cy, h, Ic);
    double2 err = P.I;
    for (int c = 0; c < 4; ++c) err = csub(err, Ic[c]);
    bool accept = hypot(err.x, err.y) < Q.eps;
    if (!accept && (P.depth + 1 >= QMAXDEPTH || sp + 4 > QSTK)) {
      accept = true;
      capped = true;
    }
    if (accept) {
      total = cadd(total, P.I);
    } else {
      if (threadIdx.x == 0)
        for (int c = 3; c >= 0; --c) stk[sp + (3 - c)] = Panel{cx[c], cy[c], h, Ic[c], P.depth + 1};
      sp += 4;
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    out[blockIdx.x] = total;
    if (capped) atomicAdd(depth_flag, 1);
  }
}

void launch_quadrature(const QuadCtx& Q, int ntask, const QuadTask* tasks, double2* out, int* depth_flag,
                       hipStream_t s) {
  if (!ntask) return;
  hipLaunchKernelGGL(k_quad, dim3(ntask), dim3(64), 0, s, Q, tasks, out, depth_flag);
  HIP_OK(hipGetLastError());
}

}  // namespace aifmm
