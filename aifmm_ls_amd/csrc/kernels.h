// Descriptor structs and launchers of the gfx950 kernels.
#pragma once
#include <vector>
#include "common.h"

namespace aifmm {

// A(rows[r], cols[c]) -> out[r + ld*c]
struct BlockDesc {
  const int64_t* rows;
  const int64_t* cols;
  double2* out;
  int nr, nc, ld;
};

// One ACA problem (see k_aca).  Primary = shorter dimension.
struct AcaProblem {
  const int64_t* rows;  // N1 point indices
  const int64_t* cols;  // N2 point indices
  int np, ns;           // primary / secondary lengths
  int prim_rows;        // 1: primary = rows (N1 < N2)
  int rmax;             // capacity of U/V/Craw
  double2* U;           // rmax x np   residual crosses
  double2* V;           // rmax x ns   normalised lines
  double2* Craw;        // rmax x np   raw crosses
  double2* Lraw;        // rmax x ns   raw lines
  double2* line;        // ns
  double2* cross;       // np
  char* remP;           // np
  char* remS;           // ns
  int* pb;              // rmax primary pivots
  int* sb;              // rmax secondary pivots
  int* rank_out;
};

void launch_entries(const EntryCtx& C, int64_t n, const int64_t* I, const int64_t* J, double2* out, hipStream_t s);
void launch_assemble(const EntryCtx& C, int nblk, const BlockDesc* d, hipStream_t s);
void launch_aca(const EntryCtx& C, int nprob, const AcaProblem* d, double tol, double tol_robust, hipStream_t s);
// Dense product y = A x (N values each, one column) with every entry as
// getMatrixEntry: leaf moments, far Graf sums per (point, leaf chunk), then the
// chunk sums in order plus the near entries.  scratch holds
// dmv_scratch(N, nleaf, nchunk) double2: the moments (NTERM x nleaf) and the
// partial sums (nchunk x N).  dmv_chunks: the default chunk count.
int dmv_chunks(int64_t N, int nleaf);
inline size_t dmv_scratch(int64_t N, int nleaf, int nchunk) {
  return (size_t)NTERM * nleaf + (size_t)nchunk * N;
}
// profile = false: no profiler events (nothing is created or allocated)
void launch_dense_matvec(const EntryCtx& C, int64_t N, const double2* x, double2* y, double2* scratch, int nchunk,
                         hipStream_t s, bool profile = true);

// ----------------------------------------------------------- quadrature
// One integral per task.  kind 0: far C(n, l) = int J_n(kr |s|) e^{-i n th_s}
// T_l(s) ds; kind 1: near V(p, l) = int H_0(kr |X_p - s|) T_l(s) ds with the
// target X_p = (ra xi_p - c) / rb of a box of radius ra centred at the origin
// seen from a source box of radius rb centred at c (NearFieldInteraction1.hpp:
// 98-128).  g indexes the geometry table.
struct QuadTask {
  int kind;  // 0: far ; 1: near
  int n;     // Bessel order (far)
  int p;     // target Chebyshev point (near)
  int l;     // basis index
  int g;     // geometry
};
struct QuadGeom {
  double kr;      // kappa * r (far: the box radius; near: the source radius rb)
  double ra, rb;  // near: target and source box radii
  double cx, cy;  // near: source box centre (target box at the origin)
};
struct QuadCtx {
  const int* basis_a;
  const int* basis_b;
  const double* cheb1d;
  const QuadGeom* geom;
  int n_cheb;
  double eps;
  const double* gl_t;  // 10 nodes
  const double* gl_w;  // 10 weights
};
void launch_quadrature(const QuadCtx& Q, int ntask, const QuadTask* tasks, double2* out, int* depth_flag,
                       hipStream_t s);

// ----------------------------------------------------------- dense batched
// op: 0 = N, 1 = conjugate transpose
template <class T>
struct GemmDesc {
  const T* A;
  const T* B;
  T* C;
  int m, n, k;
  int lda, ldb, ldc;
  int opA, opB;
  double2 alpha, beta;
};
struct TileTask {
  int g, tm, tn;
};

// Gates: a descriptor with gate set runs only when *gate == want (the
// per-block precision choice of the complex64 pivot blocks, made on the
// device; null runs always).
template <class T>
struct LUDesc {
  T* A;
  int* piv;
  int n, lda;
  const int* gate;
  int want;
};

template <class T>
struct SolveDesc {
  const T* A;  // LU factors
  const int* piv;
  T* B;
  int n, nrhs, lda, ldb;
  const int* gate;
  int want;
};

// Complex64 LU of a pivot block in LDS with partial pivoting, kept in place
// when its error estimate  rho * cond  stays at or below thr (per block),
// where rho = max|U| / max|A| is the growth factor and cond = max|u_ii| /
// min|u_ii| the condition estimate from the diagonal of U.  Otherwise the
// block (left as it was) is widened and LU-factored in FP64 into a64, in the
// same launch when its FP64 factor fits the LDS.  *flag: 0 kept in complex64,
// 1 FP64 factor in a64, 2 the FP64 factor is still to be formed (by the
// launcher's widen + LU kernels gated on flag 2).  A kept factor also writes
// its diagonal, widened, into the diagonal of a64 (where the callers read
// condition estimates); *est receives the estimate when est is set.
struct GateDesc {
  float2* A;
  int* piv;
  int* flag;
  double2* a64;
  float* est;
  float thr;
  int n, lda, ld64;
};
// Solves with such factors: B (complex64) <- A^{-1} B with the complex64 factor
// (flag 0) or the FP64 one (flag 1, the columns widened and rounded back);
// flag 2 blocks are skipped (solved by the FP64 kernels gated on flag 2).
struct MixSolveDesc {
  const float2* A32;
  const double2* A64;
  const int* piv;
  const int* flag;
  float2* B;
  int n, nrhs, lda32, ld64, ldb;
};
void launch_lu_solve_mixed(int nb, int max_cols, const MixSolveDesc* d, int nmax, hipStream_t s);
void launch_lu_gate(int nb, const GateDesc* d, int nmax, hipStream_t s);

template <class T>
struct QRDesc {
  T* A;         // m x n, overwritten with R and Householder vectors
  T* tau;       // min(m,n)
  int* perm;    // n
  typename CplxOf<T>::real* work;  // 2n: running and last exact column norms
  int m, n, lda;
  double thr;   // relative rank threshold
  int* rank_out;
  double* maxpiv_out;
  int* nz_out;  // optional: number of pivots above Eigen's zero threshold
  int early;    // 1: stop at the first pivot |R_kk| <= thr * |R_00| (rank and
                //    the first rank columns of Q only; later tau are zeroed)
  double nz_eps;  // epsilon of the zero-pivot test (0: the FP64 epsilon)
};

// B <- A^{-1} B from the column-pivoted QR of a square A (ColPivHouseholderQR
// ::solve): Q^H by the first nz reflectors, R[:nz, :nz] back-substitution,
// un-permute; unknowns past nz are 0.  scratch: n x nrhs, leading dim n.
template <class T>
struct QSolveDesc {
  const T* A;
  const T* tau;
  const int* perm;
  const int* nz;
  T* B;
  T* scratch;
  int n, nrhs, lda, ldb;
};
template <class T>
void launch_qr_solve(int nb, int max_chunks, const QSolveDesc<T>* d, hipStream_t s);

template <class T>
struct QFormDesc {
  const T* A;  // Householder vectors from QRDesc
  const T* tau;
  T* Q;        // m x k output
  int m, nref, k, lda, ldq;
};

// dst(r, c) = alpha * op(src)(r, c) (+ dst if accumulate); op: trans 0 = none,
// 1 = conjugate transpose, 2 = transpose, 3 = identity (src unused),
// 4 = zero (src unused).
template <class T>
struct CopyDesc {
  const T* src;
  T* dst;
  int nr, nc, lds, ldd;
  int trans;
  int accumulate;
  double2 alpha;
};

// dst[i] = (T) src[i], i < n (FP64 operator blocks -> working precision)
template <class T>
struct CvtDesc {
  const double2* src;
  T* dst;
  int64_t n;
  const int* gate;
  int want;
};
// max_n: the longest descriptor's n (sets the chunks per descriptor)
template <class T>
void launch_cvt(int nb, const CvtDesc<T>* d, hipStream_t s, int64_t max_n);
// dst[i] = (double2) src[i], i < n (complex64 blocks -> FP64 work copies)
struct WidenDesc {
  const float2* src;
  double2* dst;
  int64_t n;
  const int* gate;
  int want;
};
void launch_widen(int nb, const WidenDesc* d, hipStream_t s, int64_t max_n);
// the longest n of a batch of conversion descriptors
template <class D>
inline int64_t max_len(const std::vector<D>& v) {
  int64_t m = 0;
  for (const auto& d : v) m = d.n > m ? d.n : m;
  return m;
}

template <class T>
void launch_gemm(int ntiles, const TileTask* tiles, const GemmDesc<T>* d, hipStream_t s);
template <class T>
void launch_lu(int nb, const LUDesc<T>* d, hipStream_t s);
template <class T>
void launch_lu_solve(int nb, int max_chunks, const SolveDesc<T>* d, hipStream_t s);
// the same for blocks of at most nmax <= lu_lds_max(sizeof(T)) rows (larger
// descriptors are skipped): factors in LDS; the solve keeps one column per
// wave in registers, rows i and i + 64 in lane i (bitwise the results of the
// two kernels above).  lu_lds_max: the largest n whose n x n block fits the
// device's LDS per workgroup, at most 128.
int lu_lds_max(int elem_bytes);
template <class T>
void launch_lu_lds(int nb, const LUDesc<T>* d, int nmax, hipStream_t s);
template <class T>
void launch_lu_solve_lds(int nb, int max_cols, const SolveDesc<T>* d, int nmax, hipStream_t s);
// CPQR launch shapes: k_cpqr (global memory, any size) or k_cpqr_reg (one
// thread per column, the column in registers: at most CPQR_REG_ROWS rows, the
// 36 points of a leaf box, and 256 columns).
enum { CPQR_GLOBAL = 0, CPQR_REG = 1 };
constexpr int CPQR_REG_ROWS = 36;
inline int cpqr_shape(int m, int n) { return m <= CPQR_REG_ROWS && n <= 256 ? CPQR_REG : CPQR_GLOBAL; }
template <class T>
void launch_cpqr(int nb, const QRDesc<T>* d, hipStream_t s, int shape);
template <class T>
void launch_qform(int nb, const QFormDesc<T>* d, hipStream_t s);
template <class T>
void launch_copy(int nb, const CopyDesc<T>* d, hipStream_t s);
// norms into out[i], or (out null) into *(double*)d[i].dst
template <class T>
void launch_fro(int nb, const CopyDesc<T>* d, double* out, hipStream_t s);
// refinement vectors: r = b - ax with per-block partial squared norms
// (resid_blocks(n) blocks per column), and y += x over n entries
int resid_blocks(int64_t n);
void launch_resid(const double2* b, const double2* ax, double2* r, int64_t n, int nrhs, double* partial,
                  hipStream_t s);
void launch_axpy(double2* y, const double2* x, int64_t n, hipStream_t s);
// Krylov vectors (N x nrhs each, ld N; basis vector i at V + i * vstride):
// per-block partial dots conj(V_i) . w (resid_blocks(n) blocks per (i, column)),
// and w = beta_c w + sum_i coef[c * nv + i] V_i
void launch_multidot(const double2* V, int64_t vstride, int nv, const double2* w, int64_t n, int nrhs,
                     double2* partial, hipStream_t s);
void launch_lincomb(double2* w, const double2* V, int64_t vstride, int nv, const double2* coef, const double2* beta,
                    int64_t n, int nrhs, hipStream_t s);
// out[i * nv + q] = sum_b part[off[i] + q * nblk[i] + b] for n problems (real
// != 0: the nblk[i] real partials at (double*)(part + off[i]), nv = 1)
void launch_sum_partials(const double2* part, const int64_t* off, const int* nblk, int n, int nv, int real,
                         double2* out, hipStream_t s);
void launch_d2f(const double2* src, float2* dst, int64_t n, hipStream_t s);
// dst[i] = src[i] over n 16-byte elements: a plain stream for calibrating
// hardware traffic counters
void launch_stream_copy(const double2* src, double2* dst, int64_t n, hipStream_t s);
// dst[i] = *src[i], i < n (scattered device ints, e.g. per-factor pivot counts)
void launch_gather_i32(const int* const* src, int* dst, int n, hipStream_t s);
void launch_f2d(const float2* src, double2* dst, int64_t n, hipStream_t s);

// Near field of the FP64 operator, rebuilt from the translation-invariant
// tables instead of read from the assembled blocks: for every leaf box k and
// each of its neighbour slots (8 neighbours and the box itself),
// y_k += P2P(k, nn) x_nn with P2P(k, nn)[ii, jj] = kappa^2 q_i T_nb[ii, jj]
// (+ 1 on the diagonal), at rounding level from the assembled product.
struct NearDesc {
  const double2* T;  // kappa^2 x the 9 tables R x R (column-major; slot 8: the box itself)
  const double* q;   // contrast at the N points
  const int* nbr;    // nbox x 9 neighbour boxes (-1: none; slot 8: the box)
  const double2* x;  // N x nrhs, ld N
  double2* y;        // accumulated into
  double kappa;
  int L, R, nbox, nrhs;
  int64_t N;
};
// grid (box groups, members, column groups); max_nrhs = the members' largest
// column count
void launch_near(int nmem, int max_box, const NearDesc* d, hipStream_t s, int max_nrhs = 1);

// The same vector operations for several independent vectors in one launch
// (one single-column problem each; member i on grid row i).  Every member
// uses its own block count resid_blocks(n) and row stride, so its partial
// sums are bitwise those of the single-member launches.
struct VecMember {
  const double2* V;      // basis vectors, V_q = V + q * vstride
  int64_t vstride;
  double2* w;            // dot / update target (multidot, lincomb)
  const double2* b;      // resid: r = b - ax (ax may be null)
  const double2* ax;
  double2* r;            // resid output (may be null: norms only)
  int64_t n;
  double2* part;         // partial sums (multidot: complex, resid: real)
  const double2* coef;   // lincomb: nv coefficients
  const double2* beta;   // lincomb: w scale
  int nv;
};
void launch_multidot_batch(int nmem, int max_blocks, const VecMember* mem, hipStream_t s);
void launch_lincomb_batch(int nmem, int max_blocks, const VecMember* mem, hipStream_t s);
void launch_resid_batch(int nmem, int max_blocks, const VecMember* mem, hipStream_t s);

}  // namespace aifmm
