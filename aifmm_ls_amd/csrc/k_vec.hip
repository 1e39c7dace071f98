// Vector kernels of the refinement loop: residual with per-block partial
// squared norms (summed on the host in a fixed order, so the norm does not
// depend on scheduling), and y += x.
#include "cplx_dev.h"
#include "kernels.h"

namespace aifmm {

constexpr int VNT = 256;

// r = b - ax (ax may be null: r = b); r may be null (norms only).
// partial[c * gridDim.x + blockIdx.x] = sum over the block's rows of |r|^2.
__global__ __launch_bounds__(VNT) void k_resid(const double2* __restrict__ b, const double2* __restrict__ ax,
                                               double2* __restrict__ r, int64_t n, double* __restrict__ partial) {
  const int c = blockIdx.y;
  const size_t base = (size_t)n * c;
  double s = 0;
  for (int64_t i = (int64_t)blockIdx.x * VNT + threadIdx.x; i < n; i += (int64_t)gridDim.x * VNT) {
    double2 v = b[base + i];
    if (ax) {
      double2 a = ax[base + i];
      v.x -= a.x;
      v.y -= a.y;
    }
    if (r) r[base + i] = v;
    s += v.x * v.x + v.y * v.y;
  }
  for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o);
  __shared__ double sh[VNT / 64];
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0;
    for (int w = 0; w < VNT / 64; ++w) t += sh[w];
    partial[(size_t)c * gridDim.x + blockIdx.x] = t;
  }
}

__global__ __launch_bounds__(VNT) void k_axpy(double2* __restrict__ y, const double2* __restrict__ x, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * VNT + threadIdx.x; i < n; i += (int64_t)gridDim.x * VNT) {
    double2 a = y[i], b = x[i];
    a.x += b.x;
    a.y += b.y;
    y[i] = a;
  }
}

// Dots of w against nv basis vectors, per right-hand side:
// partial[((c * nv) + i) * gridDim.x + blockIdx.x] = sum over the block's rows
// of conj(V_i[n, c]) w[n, c];  V_i = V + i * vstride (each N x nrhs, ld N).
__global__ __launch_bounds__(VNT) void k_multidot(const double2* __restrict__ V, int64_t vstride, int nv,
                                                  const double2* __restrict__ w, int64_t n,
                                                  double2* __restrict__ partial) {
  const int c = blockIdx.y;
  const size_t base = (size_t)n * c;
  __shared__ double2 sh[VNT / 64];
  for (int i = 0; i < nv; ++i) {
    const double2* v = V + (size_t)i * vstride + base;
    double2 s = mk(0.0, 0.0);
    for (int64_t r = (int64_t)blockIdx.x * VNT + threadIdx.x; r < n; r += (int64_t)gridDim.x * VNT) {
      double2 a = v[r], b = w[base + r];
      s.x += a.x * b.x + a.y * b.y;
      s.y += a.x * b.y - a.y * b.x;
    }
    for (int o = 32; o > 0; o >>= 1) {
      s.x += __shfl_xor(s.x, o);
      s.y += __shfl_xor(s.y, o);
    }
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
      double2 t = mk(0.0, 0.0);
      for (int q = 0; q < VNT / 64; ++q) t.x += sh[q].x, t.y += sh[q].y;
      partial[((size_t)c * nv + i) * gridDim.x + blockIdx.x] = t;
    }
    __syncthreads();
  }
}

// w[n, c] = beta_c * w[n, c] + sum_i coef[c * nv + i] * V_i[n, c]
__global__ __launch_bounds__(VNT) void k_lincomb(double2* __restrict__ w, const double2* __restrict__ V,
                                                 int64_t vstride, int nv, const double2* __restrict__ coef,
                                                 const double2* __restrict__ beta, int64_t n) {
  const int c = blockIdx.y;
  const size_t base = (size_t)n * c;
  const double2 bt = beta[c];
  for (int64_t r = (int64_t)blockIdx.x * VNT + threadIdx.x; r < n; r += (int64_t)gridDim.x * VNT) {
    double2 acc = cmul(bt, w[base + r]);
    for (int i = 0; i < nv; ++i) {
      const double2 a = coef[(size_t)c * nv + i], v = V[(size_t)i * vstride + base + r];
      acc.x += a.x * v.x - a.y * v.y;
      acc.y += a.x * v.y + a.y * v.x;
    }
    w[base + r] = acc;
  }
}

void launch_multidot(const double2* V, int64_t vstride, int nv, const double2* w, int64_t n, int nrhs,
                     double2* partial, hipStream_t s) {
  if (!n || !nrhs || !nv) return;
  hipLaunchKernelGGL(k_multidot, dim3(resid_blocks(n), nrhs), dim3(VNT), 0, s, V, vstride, nv, w, n, partial);
  HIP_OK(hipGetLastError());
}

void launch_lincomb(double2* w, const double2* V, int64_t vstride, int nv, const double2* coef, const double2* beta,
                    int64_t n, int nrhs, hipStream_t s) {
  if (!n || !nrhs) return;
  int64_t b = (n + VNT - 1) / VNT;
  hipLaunchKernelGGL(k_lincomb, dim3((unsigned)(b < 1024 ? b : 1024), nrhs), dim3(VNT), 0, s, w, V, vstride, nv, coef,
                     beta, n);
  HIP_OK(hipGetLastError());
}

// Sums of per-block partials for several problems at once: out[i * nv + q] =
// sum over b < nblk[i] of part[off[i] + q * nblk[i] + b] (complex), or of the
// real partials ((double*)(part + off[i]))[b] when real (nv = 1).  One wave64
// per (problem, q); the order of the additions is fixed.
__global__ __launch_bounds__(64) void k_sum_partials(const double2* __restrict__ part, const int64_t* __restrict__ off,
                                                     const int* __restrict__ nblk, int nv, int real,
                                                     double2* __restrict__ out) {
  const int i = blockIdx.x, q = blockIdx.y, l = threadIdx.x;
  const int nb = nblk[i];
  double2 t = make_double2(0.0, 0.0);
  if (real) {
    const double* p = (const double*)(part + off[i]);
    for (int b = l; b < nb; b += 64) t.x += p[b];
  } else {
    const double2* p = part + off[i] + (int64_t)q * nb;
    for (int b = l; b < nb; b += 64) t.x += p[b].x, t.y += p[b].y;
  }
  for (int o = 32; o > 0; o >>= 1) {
    t.x += __shfl_xor(t.x, o);
    t.y += __shfl_xor(t.y, o);
  }
  if (l == 0) out[(int64_t)i * nv + q] = t;
}

void launch_sum_partials(const double2* part, const int64_t* off, const int* nblk, int n, int nv, int real,
                         double2* out, hipStream_t s) {
  if (!n || !nv) return;
  hipLaunchKernelGGL(k_sum_partials, dim3(n, nv), dim3(64), 0, s, part, off, nblk, nv, real, out);
  HIP_OK(hipGetLastError());
}

// ---- several members per launch (grid row = member) -------------------------

// resid_blocks on the device (the same count)
__device__ __forceinline__ int member_blocks(int64_t n) {
  int64_t b = (n + VNT - 1) / VNT;
  return (int)(b < 1024 ? (b < 1 ? 1 : b) : 1024);
}

__global__ __launch_bounds__(VNT) void k_multidot_b(const VecMember* __restrict__ mem) {
  const VecMember M = mem[blockIdx.y];
  const int nb = member_blocks(M.n);
  if ((int)blockIdx.x >= nb) return;
  __shared__ double2 sh[VNT / 64];
  for (int i = 0; i < M.nv; ++i) {
    const double2* v = M.V + (size_t)i * M.vstride;
    double2 s = mk(0.0, 0.0);
    for (int64_t r = (int64_t)blockIdx.x * VNT + threadIdx.x; r < M.n; r += (int64_t)nb * VNT) {
      double2 a = v[r], b = M.w[r];
      s.x += a.x * b.x + a.y * b.y;
      s.y += a.x * b.y - a.y * b.x;
    }
    for (int o = 32; o > 0; o >>= 1) {
      s.x += __shfl_xor(s.x, o);
      s.y += __shfl_xor(s.y, o);
    }
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
      double2 t = mk(0.0, 0.0);
      for (int q = 0; q < VNT / 64; ++q) t.x += sh[q].x, t.y += sh[q].y;
      M.part[(size_t)i * nb + blockIdx.x] = t;
    }
    __syncthreads();
  }
}

__global__ __launch_bounds__(VNT) void k_lincomb_b(const VecMember* __restrict__ mem) {
  const VecMember M = mem[blockIdx.y];
  const double2 bt = *M.beta;
  for (int64_t r = (int64_t)blockIdx.x * VNT + threadIdx.x; r < M.n; r += (int64_t)gridDim.x * VNT) {
    double2 acc = cmul(bt, M.w[r]);
    for (int i = 0; i < M.nv; ++i) {
      const double2 a = M.coef[i], v = M.V[(size_t)i * M.vstride + r];
      acc.x += a.x * v.x - a.y * v.y;
      acc.y += a.x * v.y + a.y * v.x;
    }
    M.w[r] = acc;
  }
}

__global__ __launch_bounds__(VNT) void k_resid_b(const VecMember* __restrict__ mem) {
  const VecMember M = mem[blockIdx.y];
  const int nb = member_blocks(M.n);
  if ((int)blockIdx.x >= nb) return;
  double s = 0;
  for (int64_t i = (int64_t)blockIdx.x * VNT + threadIdx.x; i < M.n; i += (int64_t)nb * VNT) {
    double2 v = M.b[i];
    if (M.ax) {
      double2 a = M.ax[i];
      v.x -= a.x;
      v.y -= a.y;
    }
    if (M.r) M.r[i] = v;
    s += v.x * v.x + v.y * v.y;
  }
  for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o);
  __shared__ double sh[VNT / 64];
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0;
    for (int w = 0; w < VNT / 64; ++w) t += sh[w];
    ((double*)M.part)[blockIdx.x] = t;
  }
}

void launch_multidot_batch(int nmem, int max_blocks, const VecMember* mem, hipStream_t s) {
  if (!nmem || max_blocks < 1) return;
  hipLaunchKernelGGL(k_multidot_b, dim3(max_blocks, nmem), dim3(VNT), 0, s, mem);
  HIP_OK(hipGetLastError());
}

void launch_lincomb_batch(int nmem, int max_blocks, const VecMember* mem, hipStream_t s) {
  if (!nmem || max_blocks < 1) return;
  hipLaunchKernelGGL(k_lincomb_b, dim3(max_blocks, nmem), dim3(VNT), 0, s, mem);
  HIP_OK(hipGetLastError());
}

void launch_resid_batch(int nmem, int max_blocks, const VecMember* mem, hipStream_t s) {
  if (!nmem || max_blocks < 1) return;
  hipLaunchKernelGGL(k_resid_b, dim3(max_blocks, nmem), dim3(VNT), 0, s, mem);
  HIP_OK(hipGetLastError());
}

int resid_blocks(int64_t n) {
  int64_t b = (n + VNT - 1) / VNT;
  return (int)(b < 1024 ? (b < 1 ? 1 : b) : 1024);
}

void launch_resid(const double2* b, const double2* ax, double2* r, int64_t n, int nrhs, double* partial,
                  hipStream_t s) {
  if (!n || !nrhs) return;
  hipLaunchKernelGGL(k_resid, dim3(resid_blocks(n), nrhs), dim3(VNT), 0, s, b, ax, r, n, partial);
  HIP_OK(hipGetLastError());
}

__global__ __launch_bounds__(VNT) void k_gather_i32(const int* const* __restrict__ src, int* __restrict__ dst, int n) {
  const int i = blockIdx.x * VNT + threadIdx.x;
  if (i < n) dst[i] = *src[i];
}

// dst[i] = src[i] over n 16-byte elements (grid-stride), for calibrating
// hardware traffic counters against a stream of known size
__global__ __launch_bounds__(VNT) void k_stream_copy(const double2* __restrict__ src, double2* __restrict__ dst,
                                                     int64_t n) {
  const int64_t st = (int64_t)gridDim.x * VNT;
  int64_t i = (int64_t)blockIdx.x * VNT + threadIdx.x;
  for (; i + 3 * st < n; i += 4 * st) {  // four loads in flight per thread
    const double2 a = src[i], b = src[i + st], c = src[i + 2 * st], d = src[i + 3 * st];
    dst[i] = a;
    dst[i + st] = b;
    dst[i + 2 * st] = c;
    dst[i + 3 * st] = d;
  }
  for (; i < n; i += st) dst[i] = src[i];
}

void launch_stream_copy(const double2* src, double2* dst, int64_t n, hipStream_t s) {
  if (n > 0) hipLaunchKernelGGL(k_stream_copy, dim3(8192), dim3(VNT), 0, s, src, dst, n);
}

void launch_gather_i32(const int* const* src, int* dst, int n, hipStream_t s) {
  if (n > 0) hipLaunchKernelGGL(k_gather_i32, dim3((n + VNT - 1) / VNT), dim3(VNT), 0, s, src, dst, n);
}

void launch_axpy(double2* y, const double2* x, int64_t n, hipStream_t s) {
  if (!n) return;
  int64_t b = (n + VNT - 1) / VNT;
  hipLaunchKernelGGL(k_axpy, dim3((unsigned)(b < 2048 ? b : 2048)), dim3(VNT), 0, s, y, x, n);
  HIP_OK(hipGetLastError());
}

}  // namespace aifmm
