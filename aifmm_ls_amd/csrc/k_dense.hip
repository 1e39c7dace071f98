// Batched small dense complex kernels for the hierarchical elimination:
// LU with partial pivoting, multi-RHS LU solves, tiled GEMM with variable
// sizes, column-pivoted Householder QR with a relative rank threshold, Q
// formation and strided block copies.  One workgroup (4 wave64) per matrix
// or per output tile; every batch is described by a device array of
// descriptors built on the host.
#include <algorithm>
#include <stdexcept>
#include <type_traits>

#include "common.h"
#include "cplx_dev.h"
#include "kernels.h"

namespace aifmm {

constexpr int NT = 256;

template <class R, int NB = NT>
__device__ void block_argmax_first(R v, int i, R* shv, int* shi, R& outv, int& outi) {
  // larger wins, ties -> lower index; i < 0 marks "none"
  for (int o = 32; o > 0; o >>= 1) {
    R ov = __shfl_down(v, o);
    int oi = __shfl_down(i, o);
    if (oi >= 0 && (i < 0 || ov > v || (ov == v && oi < i))) v = ov, i = oi;
  }
  int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (lane == 0) shv[w] = v, shi[w] = i;
  __syncthreads();
  R bv = shv[0];
  int bi = shi[0];
  for (int k = 1; k < NB / 64; ++k) {
    R ov = shv[k];
    int oi = shi[k];
    if (oi >= 0 && (bi < 0 || ov > bv || (ov == bv && oi < bi))) bv = ov, bi = oi;
  }
  __syncthreads();
  outv = bv;
  outi = bi;
}

// a gated descriptor runs only when its gate holds the wanted value
__device__ __forceinline__ bool gated_off(const int* gate, int want) { return gate && *gate != want; }

template <class R, int NB = NT>
__device__ R block_maxR(R v, R* sh) {
  for (int o = 32; o > 0; o >>= 1) {
    R u = __shfl_down(v, o);
    v = u > v ? u : v;
  }
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
  __syncthreads();
  R r = sh[0];
  for (int k = 1; k < NB / 64; ++k) r = sh[k] > r ? sh[k] : r;
  __syncthreads();
  return r;
}

template <class R, int NB = NT>
__device__ R block_sumR(R v, R* sh) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o);
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
  __syncthreads();
  R r = 0;
  for (int k = 0; k < NB / 64; ++k) r += sh[k];
  __syncthreads();
  return r;
}

// ------------------------------------------------------------------ GEMM
// C = alpha op(A) op(B) + beta C, 32x32 output tile per workgroup, K step 16.
template <class T>
__global__ __launch_bounds__(NT) void k_gemm(const TileTask* __restrict__ tiles, const GemmDesc<T>* __restrict__ ds) {
  TileTask tt = tiles[blockIdx.x];
  const GemmDesc<T> d = ds[tt.g];
  __shared__ T As[16][33];
  __shared__ T Bs[16][33];
  int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;  // 16 x 16 threads, 2x2 outputs each
  int m0 = tt.tm * 32, n0 = tt.tn * 32;
  T acc[2][2];
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b) acc[a][b] = czero<T>();
  for (int k0 = 0; k0 < d.k; k0 += 16) {
    for (int e = threadIdx.x; e < 512; e += NT) {
      int mm = e & 31, kk = e >> 5;
      int gm = m0 + mm, gk = k0 + kk;
      T v = czero<T>();
      if (gm < d.m && gk < d.k) {
        if (d.opA == 0) v = d.A[gm + (size_t)d.lda * gk];
        else v = cconj(d.A[gk + (size_t)d.lda * gm]);
      }
      As[kk][mm] = v;
      int nn = e & 31;
      int gn = n0 + nn;
      T w = czero<T>();
      if (gn < d.n && gk < d.k) {
        if (d.opB == 0) w = d.B[gk + (size_t)d.ldb * gn];
        else w = cconj(d.B[gn + (size_t)d.ldb * gk]);
      }
      Bs[kk][nn] = w;
    }
    __syncthreads();
#pragma unroll
    for (int kk = 0; kk < 16; ++kk) {
      T a0 = As[kk][tx], a1 = As[kk][tx + 16];
      T b0 = Bs[kk][ty], b1 = Bs[kk][ty + 16];
      acc[0][0] = cfma(a0, b0, acc[0][0]);
      acc[0][1] = cfma(a0, b1, acc[0][1]);
      acc[1][0] = cfma(a1, b0, acc[1][0]);
      acc[1][1] = cfma(a1, b1, acc[1][1]);
    }
    __syncthreads();
  }
  T al = from2<T>(d.alpha), be = from2<T>(d.beta);
  bool bz = d.beta.x == 0.0 && d.beta.y == 0.0;
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b) {
      int gm = m0 + tx + 16 * a, gn = n0 + ty + 16 * b;
      if (gm < d.m && gn < d.n) {
        T* c = d.C + gm + (size_t)d.ldc * gn;
        T v = cmul(al, acc[a][b]);
        if (!bz) v = cfma(be, *c, v);
        *c = v;
      }
    }
}

// ------------------------------------------------------- GEMM on MFMA (complex64)
// C = alpha op(A) op(B) + beta C with one wave64 per 32x32 output tile on the
// matrix cores: the complex product is four real products (4M split,
// Re += Ar Br - Ai Bi, Im += Ar Bi + Ai Br) of v_mfma_f32_32x32x2f32 with FP32
// accumulation.  A 32x16 tile of op(A) and a 16x32 tile of op(B) are staged
// through LDS per K step with coalesced loads (conjugate transposes applied
// while staging), double-buffered through registers.  The MFMA is fed B^T as its first operand and A^T as its
// second, so it forms C^T: a lane then holds one row of C in 16 columns, and
// the stores of a wave run down columns of the column-major C.
//   v_mfma_f32_32x32x2f32 operands: lane l gives P[l % 32][l / 32] of the
//   32x2 first operand and Q[l / 32][l % 32] of the 2x32 second; result
//   element r of lane l is D[8 (r / 4) + 4 (l / 32) + r % 4][l % 32].
typedef float f32x16 __attribute__((ext_vector_type(16)));

__global__ __launch_bounds__(64) void k_gemm_mfma_c64(const TileTask* __restrict__ tiles,
                                                      const GemmDesc<float2>* __restrict__ ds) {
  const TileTask tt = tiles[blockIdx.x];
  const GemmDesc<float2> d = ds[tt.g];
  __shared__ float2 As[16][33];  // [k][m]
  __shared__ float2 Bs[16][33];  // [k][n]
  const int l = threadIdx.x;
  const int m0 = tt.tm * 32, n0 = tt.tn * 32;
  f32x16 cr, ci;
#pragma unroll
  for (int r = 0; r < 16; ++r) cr[r] = ci[r] = 0.f;
  // K blocks of 16 are double-buffered through registers: the loads of block
  // k0 + 16 are in flight while block k0 runs on the matrix cores
  float2 ra[8], rb[8];
  auto fetch = [&](int k0) {
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const int e = l + 64 * u;
      // op(A)(m, k): contiguous along m for A, along k for A^H
      const int mm = d.opA == 0 ? (e & 31) : (e >> 4), ka = d.opA == 0 ? (e >> 5) : (e & 15);
      const int gm = m0 + mm, gka = k0 + ka;
      ra[u] = make_float2(0.f, 0.f);
      if (gm < d.m && gka < d.k)
        ra[u] = d.opA == 0 ? d.A[gm + (size_t)d.lda * gka] : cconj(d.A[gka + (size_t)d.lda * gm]);
      // op(B)(k, n): contiguous along k for B, along n for B^H
      const int nn = d.opB == 0 ? (e >> 4) : (e & 31), kb = d.opB == 0 ? (e & 15) : (e >> 5);
      const int gn = n0 + nn, gkb = k0 + kb;
      rb[u] = make_float2(0.f, 0.f);
      if (gn < d.n && gkb < d.k)
        rb[u] = d.opB == 0 ? d.B[gkb + (size_t)d.ldb * gn] : cconj(d.B[gn + (size_t)d.ldb * gkb]);
    }
  };
  if (d.k > 0) fetch(0);
  for (int k0 = 0; k0 < d.k; k0 += 16) {
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const int e = l + 64 * u;
      const int mm = d.opA == 0 ? (e & 31) : (e >> 4), ka = d.opA == 0 ? (e >> 5) : (e & 15);
      const int nn = d.opB == 0 ? (e >> 4) : (e & 31), kb = d.opB == 0 ? (e & 15) : (e >> 5);
      As[ka][mm] = ra[u];
      Bs[kb][nn] = rb[u];
    }
    __syncthreads();
    if (k0 + 16 < d.k) fetch(k0 + 16);
#pragma unroll
    for (int ks = 0; ks < 8; ++ks) {
      const float2 a = As[2 * ks + (l >> 5)][l & 31];  // A^T operand: Q[k][m]
      const float2 b = Bs[2 * ks + (l >> 5)][l & 31];  // B^T operand: P[n][k]
      cr = __builtin_amdgcn_mfma_f32_32x32x2f32(b.x, a.x, cr, 0, 0, 0);
      cr = __builtin_amdgcn_mfma_f32_32x32x2f32(-b.y, a.y, cr, 0, 0, 0);
      ci = __builtin_amdgcn_mfma_f32_32x32x2f32(b.y, a.x, ci, 0, 0, 0);
      ci = __builtin_amdgcn_mfma_f32_32x32x2f32(b.x, a.y, ci, 0, 0, 0);
    }
    __syncthreads();
  }
  const float2 al = from2<float2>(d.alpha), be = from2<float2>(d.beta);
  const bool bz = d.beta.x == 0.0 && d.beta.y == 0.0;
  const int gm = m0 + (l & 31);
#pragma unroll
  for (int r = 0; r < 16; ++r) {
    const int gn = n0 + 8 * (r >> 2) + 4 * (l >> 5) + (r & 3);
    if (gm < d.m && gn < d.n) {
      float2* c = d.C + gm + (size_t)d.ldc * gn;
      float2 v = cmul(al, make_float2(cr[r], ci[r]));
      if (!bz) v = cfma(be, *c, v);
      *c = v;
    }
  }
}

// ------------------------------------------------------------------ LU
template <class T>
__global__ __launch_bounds__(NT) void k_lu(const LUDesc<T>* __restrict__ ds) {
  using R = typename CplxOf<T>::real;
  const LUDesc<T> d = ds[blockIdx.x];
  __shared__ R shv[4];
  __shared__ int shi[4];
  if (gated_off(d.gate, d.want)) return;
  const int n = d.n, ld = d.lda;
  T* A = d.A;
  for (int k = 0; k < n; ++k) {
    R bv = -1;
    int bi = -1;
    for (int i = k + threadIdx.x; i < n; i += NT) {
      R v = cabsT(A[i + (size_t)ld * k]);
      if (bi < 0 || v > bv) bv = v, bi = i;
    }
    R pv;
    int p;
    block_argmax_first(bv, bi, shv, shi, pv, p);
    if (p < 0) p = k;
    if (p != k)
      for (int j = threadIdx.x; j < n; j += NT) {
        T t = A[k + (size_t)ld * j];
        A[k + (size_t)ld * j] = A[p + (size_t)ld * j];
        A[p + (size_t)ld * j] = t;
      }
    if (threadIdx.x == 0) d.piv[k] = p;
    __syncthreads();
    T dk = A[k + (size_t)ld * k];
    bool nz = dk.x != 0 || dk.y != 0;
    if (nz)
      for (int i = k + 1 + threadIdx.x; i < n; i += NT) A[i + (size_t)ld * k] = cdiv(A[i + (size_t)ld * k], dk);
    __syncthreads();
    int m = n - k - 1;
    if (nz && m > 0)
      for (int e = threadIdx.x; e < m * m; e += NT) {
        int i = k + 1 + e % m, j = k + 1 + e / m;
        A[i + (size_t)ld * j] = cfms(A[i + (size_t)ld * k], A[k + (size_t)ld * j], A[i + (size_t)ld * j]);
      }
    __syncthreads();
  }
}

// The same LU for a block held in LDS (S: n x n, leading dimension n), with
// the same operations in the same order (bitwise the global-memory kernel's
// result).  Every thread of the workgroup calls it.
template <class T>
__device__ void lu_in_lds(T* S, int n, int* piv) {
  using R = typename CplxOf<T>::real;
  __shared__ R shv[4];
  __shared__ int shi[4];
  for (int k = 0; k < n; ++k) {
    R bv = -1;
    int bi = -1;
    for (int i = k + threadIdx.x; i < n; i += NT) {
      R v = cabsT(S[i + n * k]);
      if (bi < 0 || v > bv) bv = v, bi = i;
    }
    R pv;
    int p;
    block_argmax_first(bv, bi, shv, shi, pv, p);
    if (p < 0) p = k;
    if (p != k)
      for (int j = threadIdx.x; j < n; j += NT) {
        T t = S[k + n * j];
        S[k + n * j] = S[p + n * j];
        S[p + n * j] = t;
      }
    if (threadIdx.x == 0) piv[k] = p;
    __syncthreads();
    T dk = S[k + n * k];
    bool nz = dk.x != 0 || dk.y != 0;
    if (nz)
      for (int i = k + 1 + threadIdx.x; i < n; i += NT) S[i + n * k] = cdiv(S[i + n * k], dk);
    __syncthreads();
    int m = n - k - 1;
    if (nz && m > 0)
      for (int e = threadIdx.x; e < m * m; e += NT) {
        int i = k + 1 + e % m, j = k + 1 + e / m;
        S[i + n * j] = cfms(S[i + n * k], S[k + n * j], S[i + n * j]);
      }
    __syncthreads();
  }
}

// LU of a block that fits the LDS (n <= nmax <= lu_lds_max): loaded once,
// factored in LDS, written back.
template <class T>
__global__ __launch_bounds__(NT) void k_lu_lds(const LUDesc<T>* __restrict__ ds, int nmax) {
  extern __shared__ __align__(16) unsigned char lu_lds[];
  T* S = (T*)lu_lds;
  const LUDesc<T> d = ds[blockIdx.x];
  const int n = d.n, ld = d.lda;
  if (n > nmax) return;  // the launcher gives such blocks to k_lu
  if (gated_off(d.gate, d.want)) return;
  for (int e = threadIdx.x; e < n * n; e += NT) S[e] = d.A[e % n + (size_t)ld * (e / n)];
  __syncthreads();
  lu_in_lds(S, n, d.piv);
  for (int e = threadIdx.x; e < n * n; e += NT) d.A[e % n + (size_t)ld * (e / n)] = S[e];
}

// Complex64 pivot-block LU with the precision chosen per block (GateDesc):
// the complex64 LU in LDS, then for a block that fails the gate the FP64 LU
// of the widened block in the same LDS (n <= n64max), or flag 2 for the
// launcher's FP64 kernels (larger blocks).  The FP64 factor is bitwise that
// of widening the block and running k_lu_lds<double2> on it.
__global__ __launch_bounds__(NT) void k_lu_gate(const GateDesc* __restrict__ ds, int nmax, int n64max) {
  extern __shared__ __align__(16) unsigned char lug_lds[];
  float2* S = (float2*)lug_lds;
  double2* S64 = (double2*)lug_lds;
  __shared__ float shr[4];
  const GateDesc d = ds[blockIdx.x];
  const int n = d.n, ld = d.lda;
  if (n > nmax) {  // the launcher gates only blocks that fit; FP64 elsewhere
    if (threadIdx.x == 0) {
      *d.flag = 2;
      if (d.est) *d.est = INFINITY;
    }
    return;
  }
  float amax = 0;
  for (int e = threadIdx.x; e < n * n; e += NT) {
    const float2 a = d.A[e % n + (size_t)ld * (e / n)];
    S[e] = a;
    amax = fmaxf(amax, hypotf(a.x, a.y));
  }
  amax = block_maxR(amax, shr);  // (its barriers also order the loads)
  lu_in_lds(S, n, d.piv);
  float umax = 0, dmax = 0, dmin = INFINITY;
  for (int e = threadIdx.x; e < n * n; e += NT) {
    const int i = e % n, j = e / n;
    if (i > j) continue;
    const float u = hypotf(S[e].x, S[e].y);
    umax = fmaxf(umax, u);
    if (i == j) dmax = fmaxf(dmax, u), dmin = fminf(dmin, u);
  }
  umax = block_maxR(umax, shr);
  dmax = block_maxR(dmax, shr);
  dmin = -block_maxR(-dmin, shr);
  // a zero or non-finite pivot gives an infinite (or NaN) estimate: FP64
  const float est = (umax / amax) * (dmax / dmin);
  const bool keep = amax > 0 && est <= d.thr;  // uniform over the workgroup
  int flag = 0;
  if (keep) {
    for (int e = threadIdx.x; e < n * n; e += NT) d.A[e % n + (size_t)ld * (e / n)] = S[e];
    for (int i = threadIdx.x; i < n; i += NT)
      d.a64[i + (size_t)d.ld64 * i] = mk((double)S[i + n * i].x, (double)S[i + n * i].y);
  } else if (n <= n64max) {
    // the block in global memory is untouched: widen it into the LDS
    for (int e = threadIdx.x; e < n * n; e += NT) {
      const float2 a = d.A[e % n + (size_t)ld * (e / n)];
      S64[e] = mk((double)a.x, (double)a.y);
    }
    __syncthreads();
    lu_in_lds(S64, n, d.piv);
    for (int e = threadIdx.x; e < n * n; e += NT) d.a64[e % n + (size_t)d.ld64 * (e / n)] = S64[e];
    flag = 1;
  } else {
    flag = 2;
  }
  if (threadIdx.x == 0) {
    *d.flag = flag;
    if (d.est) *d.est = est;
  }
}

template <class T>
__device__ __forceinline__ T shfl_c(T v, int lane) {
  v.x = __shfl(v.x, lane);
  v.y = __shfl(v.y, lane);
  return v;
}

template <class TF, class TB>
__device__ __forceinline__ TF to_f(TB b);
template <>
__device__ __forceinline__ double2 to_f<double2, double2>(double2 b) { return b; }
template <>
__device__ __forceinline__ float2 to_f<float2, float2>(float2 b) { return b; }
template <>
__device__ __forceinline__ double2 to_f<double2, float2>(float2 b) { return mk((double)b.x, (double)b.y); }
template <class TB, class TF>
__device__ __forceinline__ TB to_b(TF v);
template <>
__device__ __forceinline__ double2 to_b<double2, double2>(double2 v) { return v; }
template <>
__device__ __forceinline__ float2 to_b<float2, float2>(float2 v) { return v; }
template <>
__device__ __forceinline__ float2 to_b<float2, double2>(double2 v) { return mkf((float)v.x, (float)v.y); }

// The substitutions of nc right-hand-side columns (B, leading dimension lb;
// nc <= WS_COLS * the workgroup's waves) with the factors S (n x n in LDS,
// n <= 128) and the row order perm (in LDS): wave w takes columns w, w + 4,
// ... together (their shuffles and the factor's LDS reads overlap; each
// factor element is read once for all of them), rows i and i + 64 in lane i,
// so no workgroup barriers; arithmetic in TF, the columns stored as TB.  The
// same operations per element in the same order as k_lu_solve.
constexpr int WS_COLS = 4;
template <class TF, class TB>
__device__ void wave_solve(const TF* S, int n, const int* perm, TB* B0, int lb, int nc) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int hi = lane + 64;  // this lane's second row
  constexpr int NW = NT / 64;
  TF b0[WS_COLS], b1[WS_COLS];
#pragma unroll
  for (int u = 0; u < WS_COLS; ++u) {
    const int c = w + NW * u;
    b0[u] = b1[u] = czero<TF>();
    if (c < nc) {
      const TB* B = B0 + (size_t)lb * c;
      if (lane < n) b0[u] = to_f<TF, TB>(B[perm[lane]]);
      if (hi < n) b1[u] = to_f<TF, TB>(B[perm[hi]]);
    }
  }
  for (int k = 0; k < n - 1; ++k) {
    TF bk[WS_COLS];
#pragma unroll
    for (int u = 0; u < WS_COLS; ++u) bk[u] = shfl_c(k < 64 ? b0[u] : b1[u], k & 63);
    if (lane > k && lane < n) {
      const TF a = S[lane + n * k];
#pragma unroll
      for (int u = 0; u < WS_COLS; ++u) b0[u] = cfms(a, bk[u], b0[u]);
    }
    if (hi > k && hi < n) {
      const TF a = S[hi + n * k];
#pragma unroll
      for (int u = 0; u < WS_COLS; ++u) b1[u] = cfms(a, bk[u], b1[u]);
    }
  }
  for (int k = n - 1; k >= 0; --k) {
    const TF dk = S[k + n * k];
    const bool nzp = dk.x != 0 || dk.y != 0;
#pragma unroll
    for (int u = 0; u < WS_COLS; ++u) {
      if (k < 64) {
        if (lane == k) b0[u] = nzp ? cdiv(b0[u], dk) : czero<TF>();
      } else if (lane == k - 64) {
        b1[u] = nzp ? cdiv(b1[u], dk) : czero<TF>();
      }
    }
    TF bk[WS_COLS];
#pragma unroll
    for (int u = 0; u < WS_COLS; ++u) bk[u] = shfl_c(k < 64 ? b0[u] : b1[u], k & 63);
    if (lane < k) {
      const TF a = S[lane + n * k];
#pragma unroll
      for (int u = 0; u < WS_COLS; ++u) b0[u] = cfms(a, bk[u], b0[u]);
    }
    if (hi < k) {
      const TF a = S[hi + n * k];
#pragma unroll
      for (int u = 0; u < WS_COLS; ++u) b1[u] = cfms(a, bk[u], b1[u]);
    }
  }
#pragma unroll
  for (int u = 0; u < WS_COLS; ++u) {
    const int c = w + NW * u;
    if (c < nc) {
      TB* B = B0 + (size_t)lb * c;
      if (lane < n) B[lane] = to_b<TB, TF>(b0[u]);
      if (hi < n) B[hi] = to_b<TB, TF>(b1[u]);
    }
  }
}

// the row order of the interchanges k <-> piv[k], k = 0..n-1, applied in order
__device__ __forceinline__ void perm_from_piv(const int* piv, int n, int* perm) {
  if (threadIdx.x == 0) {
    for (int i = 0; i < n; ++i) perm[i] = i;
    for (int k = 0; k < n; ++k) {
      const int p = piv[k];
      const int t = perm[k];
      perm[k] = perm[p];
      perm[p] = t;
    }
  }
}

// LU solves of a block with n <= 128 rows, the factors in LDS (bitwise
// k_lu_solve).  grid = (matrix, group of LUS_COLS columns)
constexpr int LUS_COLS = WS_COLS * (NT / 64);
template <class T>
__global__ __launch_bounds__(NT) void k_lu_solve_wave(const SolveDesc<T>* __restrict__ ds, int nmax) {
  extern __shared__ __align__(16) unsigned char lus_lds[];
  T* S = (T*)lus_lds;
  __shared__ int perm[128];
  const SolveDesc<T> d = ds[blockIdx.x];
  const int n = d.n, la = d.lda;
  const int c0 = blockIdx.y * LUS_COLS;
  if (n > nmax || c0 >= d.nrhs) return;  // the launcher gives larger blocks to k_lu_solve
  if (gated_off(d.gate, d.want)) return;
  const int nc = d.nrhs - c0 < LUS_COLS ? d.nrhs - c0 : LUS_COLS;
  for (int e = threadIdx.x; e < n * n; e += NT) S[e] = d.A[e % n + (size_t)la * (e / n)];
  perm_from_piv(d.piv, n, perm);
  __syncthreads();
  wave_solve<T, T>(S, n, perm, d.B + (size_t)d.ldb * c0, d.ldb, nc);
}

// Solves with the gated complex64 pivot blocks (MixSolveDesc): flag 0, the
// complex64 factor; flag 1, the FP64 factor with the complex64 columns
// widened on load and rounded on store (bitwise widen + k_lu_solve_wave
// <double2> + narrow); flag 2 blocks are left to the launcher's FP64 kernels.
__global__ __launch_bounds__(NT) void k_lu_solve_mixed(const MixSolveDesc* __restrict__ ds, int nmax) {
  extern __shared__ __align__(16) unsigned char lum_lds[];
  __shared__ int perm[128];
  const MixSolveDesc d = ds[blockIdx.x];
  const int n = d.n;
  const int c0 = blockIdx.y * LUS_COLS;
  if (n > nmax || c0 >= d.nrhs) return;
  const int f = *d.flag;
  if (f == 2) return;
  const int nc = d.nrhs - c0 < LUS_COLS ? d.nrhs - c0 : LUS_COLS;
  perm_from_piv(d.piv, n, perm);
  if (f == 0) {
    float2* S = (float2*)lum_lds;
    for (int e = threadIdx.x; e < n * n; e += NT) S[e] = d.A32[e % n + (size_t)d.lda32 * (e / n)];
    __syncthreads();
    wave_solve<float2, float2>(S, n, perm, d.B + (size_t)d.ldb * c0, d.ldb, nc);
  } else {
    double2* S = (double2*)lum_lds;
    for (int e = threadIdx.x; e < n * n; e += NT) S[e] = d.A64[e % n + (size_t)d.ld64 * (e / n)];
    __syncthreads();
    wave_solve<double2, float2>(S, n, perm, d.B + (size_t)d.ldb * c0, d.ldb, nc);
  }
}

// grid = (matrix, 64-column chunk of the right-hand sides)
template <class T>
__global__ __launch_bounds__(NT) void k_lu_solve(const SolveDesc<T>* __restrict__ ds) {
  const SolveDesc<T> d = ds[blockIdx.x];
  int c0 = blockIdx.y * 64;
  if (c0 >= d.nrhs) return;
  if (gated_off(d.gate, d.want)) return;
  int nc = d.nrhs - c0 < 64 ? d.nrhs - c0 : 64;
  const int n = d.n, la = d.lda, lb = d.ldb;
  const T* A = d.A;
  T* B = d.B + (size_t)lb * c0;
  if (threadIdx.x < nc) {
    T* b = B + (size_t)lb * threadIdx.x;
    for (int k = 0; k < n; ++k) {
      int p = d.piv[k];
      if (p != k) {
        T t = b[k];
        b[k] = b[p];
        b[p] = t;
      }
    }
  }
  __syncthreads();
  for (int k = 0; k < n - 1; ++k) {
    int m = n - k - 1;
    for (int e = threadIdx.x; e < m * nc; e += NT) {
      int i = k + 1 + e % m, c = e / m;
      B[i + (size_t)lb * c] = cfms(A[i + (size_t)la * k], B[k + (size_t)lb * c], B[i + (size_t)lb * c]);
    }
    __syncthreads();
  }
  for (int k = n - 1; k >= 0; --k) {
    // an exactly zero pivot (k_lu skips its column) gives that unknown 0, as
    // the basic solution of a singular block, instead of Inf / NaN
    const T dk = A[k + (size_t)la * k];
    const bool nzp = dk.x != 0 || dk.y != 0;
    if (threadIdx.x < nc) {
      T& bk = B[k + (size_t)lb * threadIdx.x];
      bk = nzp ? cdiv(bk, dk) : T{0, 0};
    }
    __syncthreads();
    for (int e = threadIdx.x; e < k * nc; e += NT) {
      int i = e % k, c = e / k;
      B[i + (size_t)lb * c] = cfms(A[i + (size_t)la * k], B[k + (size_t)lb * c], B[i + (size_t)lb * c]);
    }
    __syncthreads();
  }
}

// ------------------------------------------------------------------ CPQR
// Column-pivoted Householder QR.  Column norms are recomputed exactly after
// each reflection.  Reflector: I - t v v^H with v_0 = 1, beta real,
// t = conj((beta - c0)/beta), beta = -sign(Re c0) ||x||.
template <class T, int NQ>
__global__ __launch_bounds__(NQ) void k_cpqr(const QRDesc<T>* __restrict__ ds) {
  // norms, reflector scalars and the rank test in double for both element
  // types: squared moduli of complex64 entries underflow below ~1e-19
  using R = typename CplxOf<T>::real;
  const QRDesc<T> d = ds[blockIdx.x];
  __shared__ double shv[NQ / 64];
  __shared__ int shi[NQ / 64];
  __shared__ double shs[NQ / 64];
  const int m = d.m, n = d.n, ld = d.lda;
  T* A = d.A;
  // column norms as LAPACK/Eigen keep them: nrm = running (downdated) norm,
  // dir = last exactly computed norm; a downdate that loses more than
  // sqrt(eps) triggers an exact recomputation (ColPivHouseholderQR)
  R* nrm = d.work;
  R* dir = d.work + n;
  const double dd_thr = sqrt((double)(sizeof(R) == 8 ? 2.220446049250313e-16 : 1.1920928955078125e-07));
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  for (int j = threadIdx.x; j < n; j += NQ) d.perm[j] = j;
  for (int j = w; j < n; j += NQ / 64) {
    double s = 0;
    for (int i = lane; i < m; i += 64) s += (double)A[i + (size_t)ld * j].x * A[i + (size_t)ld * j].x +
                                          (double)A[i + (size_t)ld * j].y * A[i + (size_t)ld * j].y;
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o);
    if (lane == 0) nrm[j] = dir[j] = (R)sqrt(s);
  }
  __syncthreads();
  int size = m < n ? m : n;
  double maxpiv = 0;
  // Eigen's "nonzero pivots": the first k whose pivot norm^2 falls below
  // (max initial norm * eps)^2 / m * (m - k), with the double epsilon for both
  // element types (a complex64 factor that dropped pivots below the float
  // epsilon would lose the badly scaled rows of low-contrast boxes)
  double maxnorm = 0;
  for (int j = threadIdx.x; j < n; j += NQ) maxnorm = fmax(maxnorm, (double)nrm[j]);
  maxnorm = block_maxR<double, NQ>(maxnorm, shs);
  const double eps_r = d.nz_eps > 0 ? d.nz_eps : 2.220446049250313e-16;
  const double th_helper = (maxnorm * eps_r) * (maxnorm * eps_r) / (m > 0 ? m : 1);
  int nonzero = size;
  int stop = size;
  for (int k = 0; k < size; ++k) {
    double bv = -1;
    int bi = -1;
    for (int j = k + threadIdx.x; j < n; j += NQ) {
      double v = nrm[j];
      if (bi < 0 || v > bv) bv = v, bi = j;
    }
    double pv;
    int big;
    block_argmax_first<double, NQ>(bv, bi, shv, shi, pv, big);
    if (big < 0) big = k;
    if (nonzero == size && pv * pv < th_helper * (m - k)) nonzero = k;
    if (big != k) {
      for (int i = threadIdx.x; i < m; i += NQ) {
        T t = A[i + (size_t)ld * k];
        A[i + (size_t)ld * k] = A[i + (size_t)ld * big];
        A[i + (size_t)ld * big] = t;
      }
      if (threadIdx.x == 0) {
        R t = nrm[k];
        nrm[k] = nrm[big];
        nrm[big] = t;
        t = dir[k];
        dir[k] = dir[big];
        dir[big] = t;
        int q = d.perm[k];
        d.perm[k] = d.perm[big];
        d.perm[big] = q;
      }
    }
    __syncthreads();
    double tl = 0;
    for (int i = k + 1 + threadIdx.x; i < m; i += NQ)
      tl += (double)A[i + (size_t)ld * k].x * A[i + (size_t)ld * k].x +
            (double)A[i + (size_t)ld * k].y * A[i + (size_t)ld * k].y;
    tl = block_sumR<double, NQ>(tl, shs);
    double c0x = A[k + (size_t)ld * k].x, c0y = A[k + (size_t)ld * k].y;
    double beta = sqrt(c0x * c0x + c0y * c0y + tl), tx, ty;
    if (d.early && k > 0 && beta <= maxpiv * d.thr) {
      // column norms do not grow under the remaining reflections: every later
      // |R_ii| is at most this one, so the rank is k
      for (int i = k + threadIdx.x; i < size; i += NQ) d.tau[i] = czero<T>();
      stop = k;
      break;
    }
    if ((tl == 0 && c0y == 0) || beta == 0) {
      // nothing to annihilate (or a column below the double range)
      tx = ty = 0;
      beta = c0x;
    } else {
      if (c0x >= 0) beta = -beta;
      double2 den = mk(c0x - beta, c0y);
      for (int i = k + 1 + threadIdx.x; i < m; i += NQ) {
        T a = A[i + (size_t)ld * k];
        double2 q = cdiv(mk((double)a.x, (double)a.y), den);
        T o;
        o.x = (R)q.x;
        o.y = (R)q.y;
        A[i + (size_t)ld * k] = o;
      }
      // t = conj((beta - c0) / beta)
      tx = (beta - c0x) / beta;
      ty = c0y / beta;
    }
    T t;
    t.x = (R)tx;
    t.y = (R)ty;
    __syncthreads();
    if (threadIdx.x == 0) {
      d.tau[k] = t;
      T bb = czero<T>();
      bb.x = (R)beta;
      A[k + (size_t)ld * k] = bb;
    }
    if (fabs(beta) > maxpiv) maxpiv = fabs(beta);
    __syncthreads();
    // apply to trailing columns, one wave per column; refresh the norms
    for (int j = k + 1 + w; j < n; j += NQ / 64) {
      T* col = A + (size_t)ld * j;
      const T* v = A + (size_t)ld * k;
      T s = czero<T>();
      for (int i = k + 1 + lane; i < m; i += 64) s = cfma(cconj(v[i]), col[i], s);
      for (int o = 32; o > 0; o >>= 1) {
        s.x += __shfl_xor(s.x, o);
        s.y += __shfl_xor(s.y, o);
      }
      s = cadd(s, col[k]);
      s = cmul(t, s);
      if (lane == 0) col[k] = csub(col[k], s);
      double nn = 0;
      for (int i = k + 1 + lane; i < m; i += 64) {
        T x = cfms(s, v[i], col[i]);
        col[i] = x;
        nn += (double)x.x * x.x + (double)x.y * x.y;
      }
      for (int o = 32; o > 0; o >>= 1) nn += __shfl_xor(nn, o);
      if (lane == 0) {
        double u = nrm[j];
        if (u != 0) {
          double tq = hypot((double)col[k].x, (double)col[k].y) / u;
          tq = (1 + tq) * (1 - tq);
          if (tq < 0) tq = 0;
          double rr = u / (double)dir[j];
          if (tq * rr * rr <= dd_thr) nrm[j] = dir[j] = (R)sqrt(nn);
          else nrm[j] = (R)(u * sqrt(tq));
        }
      }
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    int r = 0;
    double pre = maxpiv * d.thr;
    for (int i = 0; i < stop; ++i)
      if (hypot((double)A[i + (size_t)ld * i].x, (double)A[i + (size_t)ld * i].y) > pre) ++r;
    *d.rank_out = r;
    if (d.maxpiv_out) *d.maxpiv_out = maxpiv;
    if (d.nz_out) *d.nz_out = nonzero;
  }
}

// The same column-pivoted QR for a matrix of at most MR rows and NQ columns
// with one thread per column and the column in registers: pivot choice (the
// largest running norm, the first position on ties), reflectors, early stop,
// norm downdates with exact recomputation and every output as k_cpqr, but a
// step costs two workgroup barriers and reads the reflector from LDS instead
// of making several dependent round trips to global memory.
//  - Columns keep their thread; a swap exchanges positions.  The pivot column
//    of step k is final after the step and is written to A right away; the
//    others are written to their positions at the end.
//  - The reflector sits in LDS as a full column: 0 above row k and from row m
//    on, 1 at row k, v below.  Every thread then runs the same unpredicated
//    loops: s = t (v^H c), c <- c - s v, which leaves rows above k alone
//    (s = 0 for columns that take no update).
//  - Sums run down the rows in order (k_cpqr sums across lanes), and the
//    exact norm of a column is formed only when its downdate needs it.
template <class T, int MR, int NQ>
__global__ __launch_bounds__(NQ, 1) void k_cpqr_reg(const QRDesc<T>* __restrict__ ds) {
  using R = typename CplxOf<T>::real;
  const QRDesc<T> d = ds[blockIdx.x];
  __shared__ double shv[NQ / 64];
  __shared__ int shi[NQ / 64];
  __shared__ T sv[MR];          // the pivot column, then the reflector
  __shared__ double sdiag[MR];  // |R_kk| as stored
  __shared__ double ssc[6];     // beta, t, step kind
  const int m = d.m, n = d.n, ld = d.lda;
  const int j = threadIdx.x;
  const bool own = j < n;
  const double dd_thr = sqrt((double)(sizeof(R) == 8 ? 2.220446049250313e-16 : 1.1920928955078125e-07));
  // every load in flight at once: rows past m read row m - 1 (m >= 1) and
  // are zeroed, threads past n read column 0 and are zeroed
  // (real and imaginary parts as two plain arrays: they stay in registers)
  R cre[MR], cim[MR];
  auto cget = [&](int i) {
    T v;
    v.x = cre[i];
    v.y = cim[i];
    return v;
  };
  auto cset = [&](int i, T v) {
    cre[i] = v.x;
    cim[i] = v.y;
  };
  {
    const T* col = d.A + (size_t)ld * (own ? j : 0);
#pragma unroll
    for (int i = 0; i < MR; ++i) cset(i, col[i < m ? i : m - 1]);
  }
  double s0 = 0;
#pragma unroll
  for (int i = 0; i < MR; ++i) {
    if (!own || i >= m) cset(i, czero<T>());
    s0 += (double)cre[i] * cre[i] + (double)cim[i] * cim[i];
  }
  R nrm = (R)sqrt(s0), dir = nrm;
  int pos = j;
  bool done = false;  // this column was a pivot and is written
  const double maxnorm = block_maxR<double, NQ>(own ? (double)nrm : 0.0, shv);
  const double eps_r = d.nz_eps > 0 ? d.nz_eps : 2.220446049250313e-16;
  const double th_helper = (maxnorm * eps_r) * (maxnorm * eps_r) / (m > 0 ? m : 1);
  const int size = m < n ? m : n;
  int nonzero = size, stop = size;
  double maxpiv = 0;
  for (int k = 0; k < size; ++k) {
    double pv;
    int big;
    const bool cand = own && pos >= k;
    block_argmax_first<double, NQ>(cand ? (double)nrm : -1.0, cand ? pos : -1, shv, shi, pv, big);
    if (big < 0) big = k;
    if (nonzero == size && pv * pv < th_helper * (m - k)) nonzero = k;
    const bool piv = own && pos == big;
    if (own && pos == k && big != k) pos = big;
    if (piv) {
      pos = k;
#pragma unroll
      for (int i = 0; i < MR; ++i) sv[i] = cget(i);  // the raw pivot column
    }
    __syncthreads();
    // wave 0, one row per lane: tail norm, beta and the reflector's scalars
    // (every lane alike), the reflector into LDS and the final pivot column
    // (R above row k, beta, v below) into A
    if (threadIdx.x < 64) {
      const int i = threadIdx.x;
      const T x = i < MR ? sv[i] : czero<T>();
      const T c0 = sv[k];
      double tl = (i > k && i < m) ? (double)x.x * x.x + (double)x.y * x.y : 0.0;
      for (int o = 32; o > 0; o >>= 1) tl += __shfl_xor(tl, o);
      const double c0x = c0.x, c0y = c0.y;
      double beta = sqrt(c0x * c0x + c0y * c0y + tl), tx = 0, ty = 0;
      int kind;
      if (d.early && k > 0 && beta <= maxpiv * d.thr) {
        kind = 2;  // stop: the rank is k
      } else if ((tl == 0 && c0y == 0) || beta == 0) {
        kind = 1;  // nothing to annihilate (or a column below the double range)
        beta = c0x;
      } else {
        kind = 0;
        if (c0x >= 0) beta = -beta;
        tx = (beta - c0x) / beta;
        ty = c0y / beta;
      }
      if (kind != 2) {
        // v_i = x_i / (c0 - beta) below row k (the raw column when there is
        // nothing to annihilate); the reflector in LDS: 1 at row k, 0 above
        // row k and from row m on
        T v = x;
        if (kind == 0 && i > k && i < m) {
          const double2 q = cdiv(mk((double)x.x, (double)x.y), mk(c0x - beta, c0y));
          v.x = (R)q.x;
          v.y = (R)q.y;
        }
        T bb = czero<T>();
        bb.x = (R)beta;
        if (i < m) d.A[i + (size_t)ld * k] = i < k ? x : i == k ? bb : v;
        if (i < MR) {
          T w = czero<T>();
          if (i == k) w.x = (R)1;
          if (i > k && i < m) w = v;
          sv[i] = w;
        }
        if (i == 0) {
          T t;
          t.x = (R)tx;
          t.y = (R)ty;
          d.tau[k] = t;
          sdiag[k] = hypot((double)bb.x, (double)bb.y);
        }
      }
      if (i == 0) {
        ssc[0] = beta;
        ssc[1] = tx;
        ssc[2] = ty;
        ssc[5] = kind;
      }
    }
    __syncthreads();
    const double beta = ssc[0];
    const int kind = (int)ssc[5];
    if (kind == 2) {
      for (int i = k + threadIdx.x; i < size; i += NQ) d.tau[i] = czero<T>();
      stop = k;
      break;
    }
    if (piv) {
      d.perm[k] = j;
      done = true;  // its column is in A
    }
    T t;
    t.x = (R)ssc[1];
    t.y = (R)ssc[2];
    if (fabs(beta) > maxpiv) maxpiv = fabs(beta);
    // every thread: s = t (v^H c), c <- c - s v; only trailing columns keep s
    const bool act = own && pos > k;
    T sd = czero<T>(), ck = czero<T>();
    // (the scheduling fences keep a few reflector rows in registers at a
    // time, not all of them next to the column)
#pragma unroll
    for (int i = 0; i < MR; ++i) {
      ck = i == k ? cget(i) : ck;
      sd = cfma(cconj(sv[i]), cget(i), sd);
      if (i % 8 == 7) __builtin_amdgcn_sched_barrier(0);
    }
    sd = cmul(t, sd);
    if (!act) sd = czero<T>();
#pragma unroll
    for (int i = 0; i < MR; ++i) {
      cset(i, cfms(sd, sv[i], cget(i)));
      if (i % 8 == 7) __builtin_amdgcn_sched_barrier(0);
    }
    const T cknew = csub(ck, sd);
    const double u = nrm;
    if (act && u != 0) {
      double tq = hypot((double)cknew.x, (double)cknew.y) / u;
      tq = (1 + tq) * (1 - tq);
      if (tq < 0) tq = 0;
      const double rr = u / (double)dir;
      if (tq * rr * rr <= dd_thr) {
        double nn = 0;
#pragma unroll
        for (int i = 0; i < MR; ++i)
          if (i > k && i < m) nn += (double)cre[i] * cre[i] + (double)cim[i] * cim[i];
        nrm = dir = (R)sqrt(nn);
      } else {
        nrm = (R)(u * sqrt(tq));
      }
    }
  }
  if (own && !done) {
#pragma unroll
    for (int i = 0; i < MR; ++i)
      if (i < m) d.A[i + (size_t)ld * pos] = cget(i);
    d.perm[pos] = j;
  }
  if (threadIdx.x == 0) {
    int r = 0;
    const double pre = maxpiv * d.thr;
    for (int i = 0; i < stop; ++i)
      if (sdiag[i] > pre) ++r;
    *d.rank_out = r;
    if (d.maxpiv_out) *d.maxpiv_out = maxpiv;
    if (d.nz_out) *d.nz_out = nonzero;
  }
}

// ColPivHouseholderQR::solve for square A, one wave per right-hand side.
// Lane l owns rows l, l + 64, ... of its column in the scratch copy, so every
// value a lane reads it wrote itself; sums and pivots move between lanes by
// shuffles only.
template <class T>
__global__ __launch_bounds__(NT) void k_qr_solve(const QSolveDesc<T>* __restrict__ ds) {
  const QSolveDesc<T> d = ds[blockIdx.x];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int n = d.n, la = d.lda;
  const int nz = *d.nz;
  const T* A = d.A;
  for (int c = blockIdx.y * 16 + w; c < d.nrhs && c < blockIdx.y * 16 + 16; c += NT / 64) {
    T* b = d.B + (size_t)d.ldb * c;
    T* x = d.scratch + (size_t)n * c;
    for (int i = lane; i < n; i += 64) x[i] = b[i];
    // x <- Q^H x
    for (int k = 0; k < nz; ++k) {
      const T* v = A + (size_t)la * k;
      T s = czero<T>();
      for (int i = k + lane; i < n; i += 64) s = (i == k) ? cadd(s, x[i]) : cfma(cconj(v[i]), x[i], s);
      for (int o = 32; o > 0; o >>= 1) {
        s.x += __shfl_xor(s.x, o);
        s.y += __shfl_xor(s.y, o);
      }
      s = cmul(d.tau[k], s);
      for (int i = k + lane; i < n; i += 64) x[i] = (i == k) ? csub(x[i], s) : cfms(s, v[i], x[i]);
    }
    // R[:nz, :nz] y = x
    for (int i = nz - 1; i >= 0; --i) {
      T yi = czero<T>();
      if (lane == (i & 63)) {
        yi = cdiv(x[i], A[i + (size_t)la * i]);
        x[i] = yi;
      }
      yi.x = __shfl(yi.x, i & 63);
      yi.y = __shfl(yi.y, i & 63);
      for (int l = lane; l < i; l += 64) x[l] = cfms(A[l + (size_t)la * i], yi, x[l]);
    }
    // b[perm[i]] = y[i] (i < nz), 0 for the rest
    for (int i = lane; i < n; i += 64) b[d.perm[i]] = i < nz ? x[i] : czero<T>();
  }
}

// Q[:, :k] = H_0^H H_1^H ... H_{s-1}^H I[:, :k]; one wave per column.
// Column j starts at reflector j: H_h^H (h > j) leaves e_j as it is (v_h is 0
// above row h), and skipping them is exact (their updates add zeros).
template <class T>
__global__ __launch_bounds__(NT) void k_qform(const QFormDesc<T>* __restrict__ ds) {
  const QFormDesc<T> d = ds[blockIdx.x];
  const int m = d.m, lda = d.lda, ldq = d.ldq;
  int size = m < d.nref ? m : d.nref;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  for (int j = w; j < d.k; j += NT / 64) {
    T* q = d.Q + (size_t)ldq * j;
    for (int i = lane; i < m; i += 64) {
      T z = czero<T>();
      if (i == j) z.x = 1;
      q[i] = z;
    }
    for (int h = (j < size ? j : size - 1); h >= 0; --h) {
      const T* v = d.A + (size_t)lda * h;
      T s = czero<T>();
      for (int i = h + 1 + lane; i < m; i += 64) s = cfma(cconj(v[i]), q[i], s);
      for (int o = 32; o > 0; o >>= 1) {
        s.x += __shfl_xor(s.x, o);
        s.y += __shfl_xor(s.y, o);
      }
      s = cadd(s, q[h]);
      s = cmul(cconj(d.tau[h]), s);
      if (lane == 0) q[h] = csub(q[h], s);
      for (int i = h + 1 + lane; i < m; i += 64) q[i] = cfms(s, v[i], q[i]);
    }
  }
}

template <class T>
__global__ __launch_bounds__(NT) void k_copy(const CopyDesc<T>* __restrict__ ds) {
  const CopyDesc<T> d = ds[blockIdx.x];
  T al = from2<T>(d.alpha);
  int tot = d.nr * d.nc;
  for (int e = threadIdx.x + blockIdx.y * NT; e < tot; e += NT * gridDim.y) {
    int r = e % d.nr, c = e / d.nr;
    T v;
    if (d.trans == 1) v = cconj(d.src[c + (size_t)d.lds * r]);
    else if (d.trans == 2) v = d.src[c + (size_t)d.lds * r];
    else if (d.trans == 3) {
      v = czero<T>();
      if (r == c) v.x = 1;
    } else if (d.trans == 4) v = czero<T>();
    else v = d.src[r + (size_t)d.lds * c];
    v = cmul(al, v);
    T* o = d.dst + r + (size_t)d.ldd * c;
    if (d.accumulate) v = cadd(v, *o);
    *o = v;
  }
}

// Frobenius norms of a batch of blocks (one workgroup each): into out[i], or
// through descriptor i's dst when out is null.
template <class T>
__global__ __launch_bounds__(NT) void k_fro(const CopyDesc<T>* __restrict__ ds, double* __restrict__ out) {
  using R = typename CplxOf<T>::real;
  __shared__ R sh[4];
  const CopyDesc<T> d = ds[blockIdx.x];
  R s = 0;
  int tot = d.nr * d.nc;
  for (int e = threadIdx.x; e < tot; e += NT) {
    int r = e % d.nr, c = e / d.nr;
    s += cabs2(d.src[r + (size_t)d.lds * c]);
  }
  s = block_sumR(s, sh);
  // out null: each descriptor names its own output in dst
  if (threadIdx.x == 0) *(out ? out + blockIdx.x : (double*)d.dst) = sqrt((double)s);
}

// grid = (descriptor, chunk): the chunks of a descriptor stride its elements
template <class T>
__global__ __launch_bounds__(NT) void k_cvt(const CvtDesc<T>* __restrict__ ds) {
  const CvtDesc<T> d = ds[blockIdx.x];
  if (gated_off(d.gate, d.want)) return;
  for (int64_t e = threadIdx.x + (int64_t)blockIdx.y * NT; e < d.n; e += (int64_t)NT * gridDim.y)
    d.dst[e] = from2<T>(d.src[e]);
}

__global__ void k_d2f(const double2* __restrict__ s, float2* __restrict__ d, int64_t n) {
  int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (t < n) d[t] = mkf((float)s[t].x, (float)s[t].y);
}
__global__ void k_f2d(const float2* __restrict__ s, double2* __restrict__ d, int64_t n) {
  int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (t < n) d[t] = mk((double)s[t].x, (double)s[t].y);
}

// ------------------------------------------------------------ launchers
// complex64 products run on the matrix cores; complex128 on the vector ALUs
template <class T>
void launch_gemm(int ntiles, const TileTask* tiles, const GemmDesc<T>* d, hipStream_t s) {
  if (!ntiles) return;
  if constexpr (std::is_same<T, float2>::value)
    hipLaunchKernelGGL(k_gemm_mfma_c64, dim3(ntiles), dim3(64), 0, s, tiles, d);
  else
    hipLaunchKernelGGL(k_gemm<T>, dim3(ntiles), dim3(NT), 0, s, tiles, d);
  HIP_OK(hipGetLastError());
}
template <class T>
void launch_lu(int nb, const LUDesc<T>* d, hipStream_t s) {
  if (!nb) return;
  hipLaunchKernelGGL(k_lu<T>, dim3(nb), dim3(NT), 0, s, d);
  HIP_OK(hipGetLastError());
}
template <class T>
void launch_lu_solve(int nb, int max_chunks, const SolveDesc<T>* d, hipStream_t s) {
  if (!nb || !max_chunks) return;
  hipLaunchKernelGGL(k_lu_solve<T>, dim3(nb, max_chunks), dim3(NT), 0, s, d);
  HIP_OK(hipGetLastError());
}
// LDS per workgroup of the current device (160 KiB on MI355X), less room for
// the kernels' static shared arrays
static size_t lds_budget() {
  static const size_t v = [] {
    int dev = 0, b = 0;
    if (hipGetDevice(&dev) != hipSuccess || hipDeviceGetAttribute(&b, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess ||
        b <= 0)
      b = 64 * 1024;
    return (size_t)b - 2048;
  }();
  return v;
}
int lu_lds_max(int elem_bytes) {
  int n = 128;
  while ((size_t)n * n * elem_bytes > lds_budget()) --n;
  return n;
}
// dynamic LDS above the default 64 KiB must be allowed per kernel
template <class K>
static void allow_lds(K kernel, size_t bytes) {
  if (bytes > 64 * 1024) HIP_OK(hipFuncSetAttribute((const void*)kernel, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes));
}
template <class T>
void launch_lu_lds(int nb, const LUDesc<T>* d, int nmax, hipStream_t s) {
  if (!nb) return;
  if (nmax > lu_lds_max(sizeof(T))) throw std::runtime_error("LU block too large for the LDS kernel");
  const size_t lds = (size_t)nmax * nmax * sizeof(T);
  allow_lds(k_lu_lds<T>, lds);
  hipLaunchKernelGGL(k_lu_lds<T>, dim3(nb), dim3(NT), lds, s, d, nmax);
  HIP_OK(hipGetLastError());
}
template <class T>
void launch_lu_solve_lds(int nb, int max_cols, const SolveDesc<T>* d, int nmax, hipStream_t s) {
  if (!nb || !max_cols) return;
  if (nmax > lu_lds_max(sizeof(T))) throw std::runtime_error("LU block too large for the LDS kernel");
  const size_t lds = (size_t)nmax * nmax * sizeof(T);
  allow_lds(k_lu_solve_wave<T>, lds);
  hipLaunchKernelGGL(k_lu_solve_wave<T>, dim3(nb, (max_cols + LUS_COLS - 1) / LUS_COLS), dim3(NT), lds, s, d, nmax);
  HIP_OK(hipGetLastError());
}
// LDS of the gated kernels: the complex64 factor of the largest block, or the
// FP64 factor of the largest block that takes the FP64 path in LDS
static size_t gate_lds(int nmax) {
  const int n64 = std::min(nmax, lu_lds_max(sizeof(double2)));
  return std::max((size_t)nmax * nmax * sizeof(float2), (size_t)n64 * n64 * sizeof(double2)) + 16;
}
void launch_lu_gate(int nb, const GateDesc* d, int nmax, hipStream_t s) {
  if (!nb) return;
  if (nmax > lu_lds_max(sizeof(float2))) throw std::runtime_error("gated LU: block too large for the LDS kernel");
  const size_t lds = gate_lds(nmax);
  allow_lds(k_lu_gate, lds);
  hipLaunchKernelGGL(k_lu_gate, dim3(nb), dim3(NT), lds, s, d, nmax, lu_lds_max(sizeof(double2)));
  HIP_OK(hipGetLastError());
}
void launch_lu_solve_mixed(int nb, int max_cols, const MixSolveDesc* d, int nmax, hipStream_t s) {
  if (!nb || !max_cols) return;
  if (nmax > lu_lds_max(sizeof(float2))) throw std::runtime_error("gated LU solve: block too large for the LDS kernel");
  const size_t lds = gate_lds(nmax);
  allow_lds(k_lu_solve_mixed, lds);
  hipLaunchKernelGGL(k_lu_solve_mixed, dim3(nb, (max_cols + LUS_COLS - 1) / LUS_COLS), dim3(NT), lds, s, d, nmax);
  HIP_OK(hipGetLastError());
}
template <class T>
void launch_qr_solve(int nb, int max_chunks, const QSolveDesc<T>* d, hipStream_t s) {
  if (!nb || !max_chunks) return;
  hipLaunchKernelGGL(k_qr_solve<T>, dim3(nb, max_chunks), dim3(NT), 0, s, d);
  HIP_OK(hipGetLastError());
}
// 1024 threads: the trailing update of a wide stack (hundreds of columns)
// runs one wave per column, so more waves per matrix shorten every pivot step
template <class T>
void launch_cpqr(int nb, const QRDesc<T>* d, hipStream_t s, int shape) {
  if (!nb) return;
  if (shape == CPQR_REG) hipLaunchKernelGGL((k_cpqr_reg<T, CPQR_REG_ROWS, 256>), dim3(nb), dim3(256), 0, s, d);
  else hipLaunchKernelGGL((k_cpqr<T, 1024>), dim3(nb), dim3(1024), 0, s, d);
  HIP_OK(hipGetLastError());
}
template <class T>
void launch_qform(int nb, const QFormDesc<T>* d, hipStream_t s) {
  if (!nb) return;
  hipLaunchKernelGGL(k_qform<T>, dim3(nb), dim3(NT), 0, s, d);
  HIP_OK(hipGetLastError());
}
template <class T>
void launch_copy(int nb, const CopyDesc<T>* d, hipStream_t s) {
  if (!nb) return;
  hipLaunchKernelGGL(k_copy<T>, dim3(nb, 4), dim3(NT), 0, s, d);
  HIP_OK(hipGetLastError());
}
template <class T>
void launch_fro(int nb, const CopyDesc<T>* d, double* out, hipStream_t s) {
  if (!nb) return;
  hipLaunchKernelGGL(k_fro<T>, dim3(nb), dim3(NT), 0, s, d, out);
  HIP_OK(hipGetLastError());
}
// chunks per descriptor for the longest one: about 4 elements per thread
static unsigned conv_chunks(int64_t max_n) {
  const int64_t c = (max_n + 4 * NT - 1) / (4 * NT);
  return (unsigned)(c < 1 ? 1 : c > 1024 ? 1024 : c);
}
template <class T>
void launch_cvt(int nb, const CvtDesc<T>* d, hipStream_t s, int64_t max_n) {
  if (!nb) return;
  hipLaunchKernelGGL(k_cvt<T>, dim3(nb, conv_chunks(max_n)), dim3(NT), 0, s, d);
  HIP_OK(hipGetLastError());
}
__global__ __launch_bounds__(NT) void k_widen(const WidenDesc* __restrict__ ds) {
  const WidenDesc d = ds[blockIdx.x];
  if (gated_off(d.gate, d.want)) return;
  for (int64_t i = threadIdx.x + (int64_t)blockIdx.y * NT; i < d.n; i += (int64_t)NT * gridDim.y)
    d.dst[i] = mk((double)d.src[i].x, (double)d.src[i].y);
}
void launch_widen(int nb, const WidenDesc* d, hipStream_t s, int64_t max_n) {
  if (!nb) return;
  hipLaunchKernelGGL(k_widen, dim3(nb, conv_chunks(max_n)), dim3(NT), 0, s, d);
  HIP_OK(hipGetLastError());
}
void launch_d2f(const double2* src, float2* dst, int64_t n, hipStream_t s) {
  if (!n) return;
  hipLaunchKernelGGL(k_d2f, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, src, dst, n);
  HIP_OK(hipGetLastError());
}
void launch_f2d(const float2* src, double2* dst, int64_t n, hipStream_t s) {
  if (!n) return;
  hipLaunchKernelGGL(k_f2d, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, src, dst, n);
  HIP_OK(hipGetLastError());
}

#define INST(T)                                                                         \
  template void launch_gemm<T>(int, const TileTask*, const GemmDesc<T>*, hipStream_t);  \
  template void launch_lu<T>(int, const LUDesc<T>*, hipStream_t);                       \
  template void launch_lu_solve<T>(int, int, const SolveDesc<T>*, hipStream_t);         \
  template void launch_lu_lds<T>(int, const LUDesc<T>*, int, hipStream_t);               \
  template void launch_lu_solve_lds<T>(int, int, const SolveDesc<T>*, int, hipStream_t);  \
  template void launch_cpqr<T>(int, const QRDesc<T>*, hipStream_t, int);                \
  template void launch_qr_solve<T>(int, int, const QSolveDesc<T>*, hipStream_t);         \
  template void launch_qform<T>(int, const QFormDesc<T>*, hipStream_t);                 \
  template void launch_copy<T>(int, const CopyDesc<T>*, hipStream_t);                   \
  template void launch_fro<T>(int, const CopyDesc<T>*, double*, hipStream_t);            \
  template void launch_cvt<T>(int, const CvtDesc<T>*, hipStream_t, int64_t);
INST(double2)
INST(float2)

}  // namespace aifmm
