// Box-program kernel (descriptors in prog.h): one workgroup (4 wave64) per
// program.  Each output row block of up to 256 rows is spread over the
// workgroup as rp rows x ns column slices (rp = rows rounded up to 16, 32 or
// a multiple of 64; ns = 256 / rp), so a wave reads a contiguous column
// segment of the matrix per step (coalesced), and slice partial sums are
// combined through LDS.  The phase-1 vector lives in LDS; right-hand sides
// are processed CH columns at a time.
#include "cplx_dev.h"
#include "prog.h"

namespace aifmm {

constexpr int PNT = 256;
constexpr int PROG_LDS = 64 * 1024;

template <class T, int CH>
__device__ __forceinline__ bool row_block_sum(int rb, int rr, const PTerm<T>* __restrict__ terms, int t0, int t1,
                                              int c0, int nc, const T* ylds, int ldy, T* red, T (&acc)[CH],
                                              int& row) {
  const int tid = threadIdx.x;
  const int rp = rr <= 16 ? 16 : rr <= 32 ? 32 : ((rr + 63) >> 6) << 6;
  const int ns = PNT / rp;
  const int i = tid % rp, s = tid / rp;
#pragma unroll
  for (int c = 0; c < CH; ++c) acc[c] = czero<T>();
  if (s < ns && i < rr) {
    for (int t = t0; t < t1; ++t) {
      const PTerm<T> tm = terms[t];
      const T* A = tm.A + rb + i;
      const T* x;
      size_t ldx;
      if (tm.x) {
        x = tm.x + (size_t)tm.ldx * c0;
        ldx = tm.ldx;
      } else {
        x = ylds;
        ldx = ldy;
      }
      T pa[CH];
#pragma unroll
      for (int c = 0; c < CH; ++c) pa[c] = czero<T>();
      for (int j = s; j < tm.cols; j += ns) {
        const T a = A[(size_t)tm.lda * j];
#pragma unroll
        for (int c = 0; c < CH; ++c)
          if (c < nc) pa[c] = cfma(a, x[j + ldx * c], pa[c]);
      }
      const T al = from2<T>(tm.alpha);
#pragma unroll
      for (int c = 0; c < CH; ++c) acc[c] = cfma(al, pa[c], acc[c]);
    }
  }
  if (ns > 1) {
#pragma unroll
    for (int c = 0; c < CH; ++c) red[tid + PNT * c] = acc[c];
    __syncthreads();
    if (s == 0 && i < rr)
      for (int q = 1; q < ns; ++q)
#pragma unroll
        for (int c = 0; c < CH; ++c) acc[c] = cadd(acc[c], red[i + q * rp + PNT * c]);
    __syncthreads();
  }
  row = rb + i;
  return s == 0 && i < rr;
}

template <class T, int CH>
__global__ __launch_bounds__(PNT) void k_prog(const PProg<T>* __restrict__ progs, const PTerm<T>* __restrict__ terms,
                                              int nrhs, int ldy) {
  extern __shared__ __align__(16) unsigned char smem[];
  T* ylds = (T*)smem;
  T* red = ylds + (size_t)ldy * CH;
  const PProg<T>* P = progs + blockIdx.x;
  const int m = P->m;
  for (int c0 = 0; c0 < nrhs; c0 += CH) {
    const int nc = min(CH, nrhs - c0);
    for (int rb = 0; rb < m; rb += PNT) {
      const int rr = min(PNT, m - rb);
      T acc[CH];
      int row;
      if (row_block_sum<T, CH>(rb, rr, terms, P->t0, P->t1, c0, nc, ylds, ldy, red, acc, row)) {
        const T sa = from2<T>(P->src_alpha);
#pragma unroll
        for (int c = 0; c < CH; ++c) {
          if (c >= nc) break;
          T v = acc[c];
          if (P->src) v = cfma(sa, P->src[row + (size_t)P->ldsrc * (c0 + c)], v);
          ylds[row + (size_t)ldy * c] = v;
          if (P->save) P->save[row + (size_t)P->ldsave * (c0 + c)] = v;
        }
      }
    }
    __syncthreads();
    for (int o = 0; o < P->nout; ++o) {
      const int rows = P->rows[o];
      for (int rb = 0; rb < rows; rb += PNT) {
        const int rr = min(PNT, rows - rb);
        T acc[CH];
        int row;
        if (row_block_sum<T, CH>(rb, rr, terms, P->u0[o], P->u1[o], c0, nc, ylds, ldy, red, acc, row)) {
          T* out = P->out[o];
          const size_t ldo = P->ldo[o];
#pragma unroll
          for (int c = 0; c < CH; ++c) {
            if (c >= nc) break;
            out[row + ldo * (c0 + c)] = acc[c];
          }
        }
      }
    }
    __syncthreads();
  }
}

static int pick_ch(int ldy, int eb) {
  for (int ch : {4, 2, 1})
    if ((size_t)(ldy + PNT) * ch * eb <= (size_t)PROG_LDS) return ch;
  return 0;
}

int prog_max_m(int elem_bytes) { return PROG_LDS / elem_bytes - PNT; }

template <class T>
void launch_prog(int nprog, const PProg<T>* progs, const PTerm<T>* terms, int nrhs, int ldy, hipStream_t s) {
  if (!nprog || !nrhs) return;
  if (ldy < 1) ldy = 1;
  const int ch = pick_ch(ldy, sizeof(T));
  if (!ch) throw std::runtime_error("box program vector of " + std::to_string(ldy) + " rows exceeds LDS");
  const size_t lds = (size_t)(ldy + PNT) * ch * sizeof(T);
  if (ch == 4) hipLaunchKernelGGL((k_prog<T, 4>), dim3(nprog), dim3(PNT), lds, s, progs, terms, nrhs, ldy);
  else if (ch == 2) hipLaunchKernelGGL((k_prog<T, 2>), dim3(nprog), dim3(PNT), lds, s, progs, terms, nrhs, ldy);
  else hipLaunchKernelGGL((k_prog<T, 1>), dim3(nprog), dim3(PNT), lds, s, progs, terms, nrhs, ldy);
  HIP_OK(hipGetLastError());
}

template void launch_prog<double2>(int, const PProg<double2>*, const PTerm<double2>*, int, int, hipStream_t);
template void launch_prog<float2>(int, const PProg<float2>*, const PTerm<float2>*, int, int, hipStream_t);

}  // namespace aifmm
