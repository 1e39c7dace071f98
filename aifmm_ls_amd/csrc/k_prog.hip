// Box-program kernel (descriptors in prog.h): one workgroup (4, 8 or 16
// wave64) per program and group of CH right-hand-side columns (blockIdx.y).
// Matrices are complex128 or complex64 (T); vectors, products and sums are
// complex128 (V).
// Each output row block of up to 256 rows is spread over the
// workgroup as rp rows x ns column slices (rp = rows rounded up to 16, 32 or
// a multiple of 64; ns = threads / rp), so a wave reads a contiguous column
// segment of the matrix per step (coalesced), and slice partial sums are
// combined through LDS.  The phase-1 vector lives in LDS; right-hand sides
// are processed CH columns at a time.  Where the launch gives room for it
// (xs_cols), the terms' input vectors of the CH columns are staged in LDS
// by coalesced loads before a sum walks them, instead of every thread
// loading each element from global memory; the values and the order of the
// arithmetic are the same either way.
#include <cstddef>

#include "cplx_dev.h"
#include "prog.h"

namespace aifmm {

using V = double2;
__device__ __forceinline__ V wv(float2 a) { return mk((double)a.x, (double)a.y); }
__device__ __forceinline__ V wv(double2 a) { return a; }

// dynamic LDS of a program (the phase-1 vector and the reduction buffer); the
// staged term descriptors take the rest of the 64 KB a workgroup may address
constexpr int PROG_LDS = 63 * 1024;

// Sum over terms of alpha_t A_t[rb + i, :] x_t for this thread's row i and
// column slice s (columns s, s + ns, ...), four independent loads in flight.
template <class T, int CH, int NTH>
__device__ __forceinline__ bool row_block_sum(int rb, int rr, const PTerm<T>* __restrict__ terms, int t0, int t1,
                                              int c0, int nc, const V* ylds, int ldy, V* red, V (&acc)[CH],
                                              int& row, V* xs, int xs_cols) {
  const int tid = threadIdx.x;
  const int rp = rr <= 16 ? 16 : rr <= 32 ? 32 : ((rr + 63) >> 6) << 6;
  const int ns = NTH / rp;
  const int i = tid % rp, s = tid / rp;
#pragma unroll
  for (int c = 0; c < CH; ++c) acc[c] = czero<V>();
  const int nt = t1 - t0;
  // One column per pass.  Complex64 (the sweeps) walks the terms' columns as
  // one sequence (path 2).  Complex128 (the FP64 operator products: many
  // terms of a few columns each, bandwidth-bound) shares the slices out among
  // the terms when there are no more terms than slices (path 1), else takes
  // the general path; both hold fewer registers than the walk.  Measured on
  // MI355X at N = 147,456: the walk takes the complex64 sweep from 1.82 to
  // 1.72 ms, and would take the complex128 operator product from 1.75 to
  // 3.2 ms per refined solve.
  const int path = CH != 1 || nt <= 0 ? 0 : sizeof(T) == 8 ? 2 : ns >= nt ? 1 : 0;
  // staged input vectors: term t's block (cols x nc, leading dimension cols)
  // at the running offset of the earlier terms' cols x CH
  const bool staged = CH > 1 && xs_cols > 0 && nt > 0;
  if (staged) {
    __syncthreads();  // the previous sum's readers are done with xs
    int off = 0;
    for (int t = t0; t < t1; ++t) {
      const PTerm<T>& tm = terms[t];
      if (!tm.x) continue;
      const int cols = tm.cols, n = cols * nc;
      const double2* x = tm.x + (size_t)tm.ldx * c0;
      for (int e = tid; e < n; e += NTH) {
        const int c = e / cols, j = e - c * cols;
        xs[off + e] = x[j + (size_t)tm.ldx * c];
      }
      off += cols * CH;
    }
    __syncthreads();
  }  if (path == 1) {
    // no more terms than slices: the slices are shared out among the terms
    // (spt each), so all terms stream at once
    const int spt = ns / nt;
    if (s < spt * nt && i < rr) {
      const PTerm<T> tm = terms[t0 + s % nt];
      const int q = s / nt;
      const T* A = tm.A + rb + i;
      const V* x = tm.x ? tm.x + (size_t)tm.ldx * c0 : ylds;
      const size_t la = tm.lda;
      // two accumulators: alternate products do not wait on each other
      V pa = czero<V>(), pb = czero<V>();
      int j = q;
      for (; j + 3 * spt < tm.cols; j += 4 * spt) {
        const T a0 = A[la * j], a1 = A[la * (j + spt)], a2 = A[la * (j + 2 * spt)], a3 = A[la * (j + 3 * spt)];
        const V x0 = x[j], x1 = x[j + spt], x2 = x[j + 2 * spt], x3 = x[j + 3 * spt];
        pa = cfma(wv(a0), x0, pa);
        pb = cfma(wv(a1), x1, pb);
        pa = cfma(wv(a2), x2, pa);
        pb = cfma(wv(a3), x3, pb);
      }
      for (; j < tm.cols; j += spt) pa = cfma(wv(A[la * j]), x[j], pa);
      pa = cadd(pa, pb);
      acc[0] = cmul(tm.alpha, pa);
    }
  } else if (sizeof(T) == 8 && path == 2) {
    // the terms' columns form one sequence (term t0's columns, then t0 + 1's,
    // ...), and slice s takes elements s, s + ns, s + 2 ns, ... of it across
    // term boundaries.  U loads are issued before any is used, so a program
    // of many terms keeps U loads in flight per thread instead of waiting at
    // every term.
    constexpr int U = 8;
    if (s < ns && i < rr) {
      // the current term's fields stay in registers; they are read again
      // (all at once) when the walk enters the next term
      const T* Ab;
      const V* xb;
      size_t la;
      V alp;
      int ct;
      int t = t0, j = s;
      auto enter = [&](int tt) {
        const PTerm<T>& tm = terms[tt];
        Ab = tm.A + rb + i;
        la = tm.lda;
        xb = tm.x ? tm.x + (size_t)tm.ldx * c0 : ylds;  // (CH = 1 here: never staged)
        alp = tm.alpha;
        ct = tm.cols;
      };
      auto land = [&]() {
        while (j >= ct) {
          j -= ct;
          if (++t == t1) return;
          enter(t);
        }
      };
      enter(t0);
      land();
      while (t < t1) {
        T av[U];
        V xv[U], al[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
          av[u] = czero<T>();
          xv[u] = al[u] = czero<V>();
          if (t < t1) {
            av[u] = Ab[la * j];
            xv[u] = xb[j];
            al[u] = alp;
            j += ns;
            land();
          }
        }
#pragma unroll
        for (int u = 0; u < U; ++u) acc[0] = cfma(al[u], cmul(wv(av[u]), xv[u]), acc[0]);
      }
    }
  } else if (s < ns && i < rr) {
    int xoff = 0;
    for (int t = t0; t < t1; ++t) {
      const PTerm<T> tm = terms[t];
      const T* A = tm.A + rb + i;
      const V* x;
      size_t ldx;
      if (tm.x && staged) {
        x = xs + xoff;
        ldx = tm.cols;
        xoff += tm.cols * CH;
      } else if (tm.x) {
        x = tm.x + (size_t)tm.ldx * c0;
        ldx = tm.ldx;
      } else {
        x = ylds;
        ldx = ldy;
      }
      const size_t la = tm.lda;
      V pa[CH];
#pragma unroll
      for (int c = 0; c < CH; ++c) pa[c] = czero<V>();
      int j = s;
      for (; j + 3 * ns < tm.cols; j += 4 * ns) {
        const T a0 = A[la * j], a1 = A[la * (j + ns)], a2 = A[la * (j + 2 * ns)], a3 = A[la * (j + 3 * ns)];
#pragma unroll
        for (int c = 0; c < CH; ++c)
          if (c < nc) {
            const V* xc = x + ldx * c;
            pa[c] = cfma(wv(a0), xc[j], pa[c]);
            pa[c] = cfma(wv(a1), xc[j + ns], pa[c]);
            pa[c] = cfma(wv(a2), xc[j + 2 * ns], pa[c]);
            pa[c] = cfma(wv(a3), xc[j + 3 * ns], pa[c]);
          }
      }
      for (; j < tm.cols; j += ns) {
        const V a = wv(A[la * j]);
#pragma unroll
        for (int c = 0; c < CH; ++c)
          if (c < nc) pa[c] = cfma(a, x[j + ldx * c], pa[c]);
      }
      const V al = tm.alpha;
#pragma unroll
      for (int c = 0; c < CH; ++c) acc[c] = cfma(al, pa[c], acc[c]);
    }
  }
  if (ns > 1 && rp < 64) {
    // rp = 16 or 32: the 64 / rp slices inside a wave hold the same rows
    // (lane = slice * rp + row); add them across lanes, then combine the
    // per-wave partials of each row in wave 0 the same way
    constexpr int NW = NTH / 64;
    const int lane = tid & 63, w = tid >> 6;
    for (int off = rp; off < 64; off <<= 1)
#pragma unroll
      for (int c = 0; c < CH; ++c) {
        acc[c].x += __shfl_xor(acc[c].x, off);
        acc[c].y += __shfl_xor(acc[c].y, off);
      }
    if (lane < rp)
#pragma unroll
      for (int c = 0; c < CH; ++c) red[w * rp + lane + NTH * c] = acc[c];
    __syncthreads();
    if (w == 0) {
      const int g = lane / rp, ng = 64 / rp;
#pragma unroll
      for (int c = 0; c < CH; ++c) acc[c] = czero<V>();
      for (int q = g; q < NW; q += ng)
#pragma unroll
        for (int c = 0; c < CH; ++c) acc[c] = cadd(acc[c], red[q * rp + i + NTH * c]);
      for (int off = rp; off < 64; off <<= 1)
#pragma unroll
        for (int c = 0; c < CH; ++c) {
          acc[c].x += __shfl_xor(acc[c].x, off);
          acc[c].y += __shfl_xor(acc[c].y, off);
        }
    }
    __syncthreads();
  } else if (ns > 1) {
#pragma unroll
    for (int c = 0; c < CH; ++c) red[tid + NTH * c] = acc[c];
    __syncthreads();
    if (s == 0 && i < rr)
      for (int q = 1; q < ns; ++q)
#pragma unroll
        for (int c = 0; c < CH; ++c) acc[c] = cadd(acc[c], red[i + q * rp + NTH * c]);
    __syncthreads();
  }
  row = rb + i;
  return s == 0 && i < rr;
}

// One box program run by the whole workgroup (ylds: ldy x CH phase-1
// vector, red: NTH x CH reduction buffer, both in LDS).
template <class T, int CH, int NTH>
__device__ __forceinline__ void run_program(const PProg<T>* __restrict__ P, const PTerm<T>* __restrict__ terms,
                                            int nrhs, int ldy, V* ylds, V* red, V* xs, int xs_cols) {
  const int m = P->m;
  // column groups of CH right-hand sides are independent (each has its own
  // phase-1 vector and writes its own columns): workgroup (program, g) runs
  // groups g, g + gridDim.y, ...
  for (int c0 = blockIdx.y * CH; c0 < nrhs; c0 += CH * gridDim.y) {
    const int nc = min(CH, nrhs - c0);
    for (int rb = 0; rb < m; rb += 256) {
      const int rr = min(256, m - rb);
      V acc[CH];
      int row;
      // the source rows are loaded ahead of the term sums (same row mapping
      // as row_block_sum: row rb + tid % rp, written by slice 0)
      V sv[CH];
      {
        const int rp = rr <= 16 ? 16 : rr <= 32 ? 32 : ((rr + 63) >> 6) << 6;
        const int i = threadIdx.x % rp;
#pragma unroll
        for (int c = 0; c < CH; ++c)
          sv[c] = (P->src && threadIdx.x < rp && i < rr && c < nc) ? P->src[rb + i + (size_t)P->ldsrc * (c0 + c)]
                                                                  : czero<V>();
      }
      if (row_block_sum<T, CH, NTH>(rb, rr, terms, P->t0, P->t1, c0, nc, ylds, ldy, red, acc, row, xs, xs_cols)) {
        const V sa = P->src_alpha;
#pragma unroll
        for (int c = 0; c < CH; ++c) {
          if (c >= nc) break;
          V v = acc[c];
          if (P->src) v = cfma(sa, sv[c], v);
          ylds[row + (size_t)ldy * c] = v;
          if (P->save) P->save[row + (size_t)P->ldsave * (c0 + c)] = v;
        }
      }
    }
    __syncthreads();
    for (int o = 0; o < P->nout; ++o) {
      const int rows = P->rows[o];
      for (int rb = 0; rb < rows; rb += 256) {
        const int rr = min(256, rows - rb);
        V acc[CH];
        int row;
        if (row_block_sum<T, CH, NTH>(rb, rr, terms, P->u0[o], P->u1[o], c0, nc, ylds, ldy, red, acc, row, xs,
                                      xs_cols)) {
          V* out = P->out[o];
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

// The program's header is read with scalar loads from its slot; its term
// descriptors are staged in LDS in one round of 16-byte loads (a program with
// more terms than a slot holds reads them from the stage-wide array).
template <class T, int CH, int NTH>
__global__ __launch_bounds__(NTH) void k_prog(const PSlot<T>* __restrict__ slots, const PTerm<T>* __restrict__ terms,
                                              int nrhs, int ldy, int slot_words, int xs_cols) {
  extern __shared__ __align__(16) unsigned char smem[];
  __shared__ PTerm<T> tl[PSLOT_TERMS];
  static_assert(sizeof(PSlot<T>) % 16 == 0 && offsetof(PSlot<T>, t) % 16 == 0, "slot layout");
  V* ylds = (V*)smem;
  V* red = ylds + (size_t)ldy * CH;
  V* xs = red + (size_t)NTH * CH;
  const PSlot<T>* S = slots + blockIdx.x;
  const int tw = slot_words - (int)(offsetof(PSlot<T>, t) / 16);
  const int4* src = (const int4*)S->t;
  int4* dst = (int4*)tl;
  for (int i = threadIdx.x; i < tw; i += NTH) dst[i] = src[i];
  __syncthreads();
  if (S->p.ext) run_program<T, CH, NTH>(&S->p, terms, nrhs, ldy, ylds, red, xs, xs_cols);
  else run_program<T, CH, NTH>(&S->p, tl, nrhs, ldy, ylds, red, xs, xs_cols);
}

// right-hand-side columns per pass: 1 for a single RHS, else the most that
// fit the LDS with the reduction buffer of NTH threads
static int pick_ch(int nrhs, int ldy, int eb, int nth) {
  if (nrhs == 1) return (size_t)(ldy + nth) * eb <= (size_t)PROG_LDS ? 1 : 0;
  for (int ch : {4, 2, 1})
    if ((size_t)(ldy + nth) * ch * eb <= (size_t)PROG_LDS) return ch;
  return 0;
}

int prog_max_m() { return PROG_LDS / (int)sizeof(V) - 1024; }

// columns per pass asked for (ch_req), checked against the LDS: 1, 2, 4, or
// 8 at up to 512 threads (8 columns of accumulators need the registers
// that 1024-thread workgroups do not leave); 0 if it does not fit
static int fit_ch(int nrhs, int ldy, int eb, int nth, int ch_req) {
  if (!ch_req) return pick_ch(nrhs, ldy, eb, nth);
  const bool known = ch_req == 1 || ch_req == 2 || ch_req == 4 || (ch_req == 8 && nth <= 512);
  if (!known || (nrhs == 1 && ch_req != 1)) return 0;
  return (size_t)(ldy + nth) * ch_req * eb <= (size_t)PROG_LDS ? ch_req : 0;
}

template <class T, int NTH>
static void launch_nth(int nprog, const PSlot<T>* progs, const PTerm<T>* terms, int nrhs, int ldy, int sw,
                       hipStream_t s, int ch_req, int xcols) {
  if (sw < 1 || sw > (int)(sizeof(PSlot<T>) / 16)) throw std::runtime_error("bad program slot size");
  const int ch = fit_ch(nrhs, ldy, sizeof(V), NTH, ch_req);
  if (!ch) throw std::runtime_error("box program vector of " + std::to_string(ldy) + " rows exceeds LDS");
  // the staged input vectors, for complex64 passes of four or more columns
  // in workgroups of up to 512 threads, where they fit beside the phase-1
  // vector and the reduction buffer.  Measured on MI355X (fixed default
  // shapes): complex64, 4 columns, 256 threads: 48.1 -> 33.6 ms over a
  // 64-column refined solve at N = 147,456 (the 256-column direct sweep at
  // N = 589,824: 212 -> 179 ms); but complex64, 2 columns, 1024 threads
  // 359 -> 395 ms, complex128, 2 columns, 1024 threads 56 -> 79 ms,
  // complex128, 4 columns, 256 threads 57.7 -> 58.7 ms, and one column
  // 0.73 -> 1.06 ms per sweep (the copy and its two barriers outweigh the
  // loads saved)
  const bool stage = sizeof(T) == 8 && ch >= 4 && NTH <= 512;
  const int xs =
      stage && xcols > 0 && (size_t)(ldy + NTH + xcols) * ch * sizeof(V) <= (size_t)PROG_LDS ? xcols : 0;
  const size_t lds = (size_t)(ldy + NTH + xs) * ch * sizeof(V);
  // one workgroup per program and group of ch right-hand-side columns
  const int groups = (nrhs + ch - 1) / ch;
  if (groups > 65535) throw std::runtime_error("too many right-hand-side column groups");
  const dim3 grid(nprog, groups);
  if constexpr (NTH <= 512) {
    if (ch == 8) {
      hipLaunchKernelGGL((k_prog<T, 8, NTH>), grid, dim3(NTH), lds, s, progs, terms, nrhs, ldy, sw, xs);
      HIP_OK(hipGetLastError());
      return;
    }
  }
  if (ch == 4) hipLaunchKernelGGL((k_prog<T, 4, NTH>), grid, dim3(NTH), lds, s, progs, terms, nrhs, ldy, sw, xs);
  else if (ch == 2)
    hipLaunchKernelGGL((k_prog<T, 2, NTH>), grid, dim3(NTH), lds, s, progs, terms, nrhs, ldy, sw, xs);
  else hipLaunchKernelGGL((k_prog<T, 1, NTH>), grid, dim3(NTH), lds, s, progs, terms, nrhs, ldy, sw, xs);
  HIP_OK(hipGetLastError());
}

// Launch shape of a stage: workgroup threads nth = 256, 512 or 1024 plus
// PROG_SHAPE_CH x right-hand-side columns per pass (0 columns: the most
// that fit the LDS, up to 4), or 0 for the default, 1024 threads, or 256 for
// stages of more than 1024 programs (those already fill the chip).  The
// program's tuner picks other shapes per stage where they measure faster.
int prog_default_nth(int nprog) { return nprog > 1024 ? 256 : 1024; }

template <class T>
bool prog_shape_fits(int nrhs, int ldy, int shape) {
  const int nth = shape % PROG_SHAPE_CH, ch = shape / PROG_SHAPE_CH;
  return (nth == 256 || nth == 512 || nth == 1024) && fit_ch(nrhs, ldy < 1 ? 1 : ldy, sizeof(V), nth, ch) > 0;
}

template <class T>
void launch_prog(int nprog, const PSlot<T>* progs, const PTerm<T>* terms, int nrhs, int ldy, int slot_words,
                 hipStream_t s, int shape, int xcols) {
  if (!nprog || !nrhs) return;
  if (ldy < 1) ldy = 1;
  int nth = shape % PROG_SHAPE_CH;
  const int ch = shape / PROG_SHAPE_CH;
  if (!nth) nth = prog_default_nth(nprog);
  if (nth == 256) launch_nth<T, 256>(nprog, progs, terms, nrhs, ldy, slot_words, s, ch, xcols);
  else if (nth == 512) launch_nth<T, 512>(nprog, progs, terms, nrhs, ldy, slot_words, s, ch, xcols);
  else if (nth == 1024) launch_nth<T, 1024>(nprog, progs, terms, nrhs, ldy, slot_words, s, ch, xcols);
  else throw std::runtime_error("box program workgroup size must be 256, 512 or 1024");
}

template void launch_prog<double2>(int, const PSlot<double2>*, const PTerm<double2>*, int, int, int, hipStream_t, int,
                                   int);
template void launch_prog<float2>(int, const PSlot<float2>*, const PTerm<float2>*, int, int, int, hipStream_t, int,
                                  int);
template bool prog_shape_fits<double2>(int, int, int);
template bool prog_shape_fits<float2>(int, int, int);

}  // namespace aifmm
