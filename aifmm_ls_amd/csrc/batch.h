// Host-side helpers that collect descriptors for the batched kernels,
// upload them and launch.
#pragma once
#include <algorithm>
#include <type_traits>
#include <vector>

#include "arena.h"
#include "kernels.h"
#include "merge.h"
#include "prof.h"

namespace aifmm {

// kernel names for the profiler: string literals, one per element type
template <class T>
constexpr bool IS_C64 = std::is_same<T, float2>::value;
#define KNAME(T, base) (IS_C64<T> ? base "<c64>" : base "<c128>")

template <class T>
struct Blk {
  T* p = nullptr;
  int r = 0, c = 0, ld = 0;
  size_t size() const { return (size_t)r * c; }
  bool empty() const { return r == 0 || c == 0; }
  Blk sub(int i0, int j0, int nr, int nc) const {
    Blk b;
    b.p = p ? p + i0 + (size_t)ld * j0 : nullptr;
    b.r = nr;
    b.c = nc;
    b.ld = ld;
    return b;
  }
};

template <class T>
Blk<T> new_blk(Arena& a, int r, int c) {
  Blk<T> b;
  b.r = r;
  b.c = c;
  b.ld = r > 0 ? r : 1;
  if ((size_t)r * c > 0) b.p = a.get<T>((size_t)r * c);
  return b;
}

// Host <-> device copies through a per-thread ring of pinned memory, ordered
// on stream s.  h2d returns once the source has been copied into the ring
// (the source may be released then); d2h returns when the data is in dst and
// synchronises stream s only.  Neither waits for other streams.
void h2d(void* dst, const void* src, size_t bytes, hipStream_t s);
// two host arrays in one copy: a at dst, b at dst + na rounded up to 256 bytes
void h2d2(void* dst, const void* a, size_t na, const void* b, size_t nb, hipStream_t s);
void d2h(void* dst, const void* src, size_t bytes, hipStream_t s);
// Streams of the library.  The staging rings and the profiler keep events
// recorded on them, and a HIP event refers to the stream it was last recorded
// on, so streams are never destroyed: stream_destroy synchronises the stream,
// folds pending profiler events and returns it to a pool that stream_create
// draws from.
hipStream_t stream_create(unsigned flags);
void stream_destroy(hipStream_t s);

// Upload a host vector into a transient device buffer in the given scratch
// arena (stream-ordered; the host vector may be released on return).
template <class D>
D* upload(Arena& a, const std::vector<D>& v, hipStream_t s) {
  if (v.empty()) return nullptr;
  const size_t bytes = v.size() * sizeof(D);
  D* d = (D*)a.alloc(bytes);
  h2d(d, v.data(), bytes, s);
  return d;
}

// The descriptors of n batches of one type as one vector (the batches are
// left empty): a merged launch of several host threads' batches.
template <class B, class D>
std::vector<D> take_all(void* const* bs, int n, std::vector<D> B::*mem) {
  std::vector<D> all;
  if (n == 1) {
    all.swap(((B*)bs[0])->*mem);
    return all;
  }
  size_t tot = 0;
  for (int i = 0; i < n; ++i) tot += (((B*)bs[i])->*mem).size();
  all.reserve(tot);
  for (int i = 0; i < n; ++i) {
    std::vector<D>& v = ((B*)bs[i])->*mem;
    all.insert(all.end(), v.begin(), v.end());
    v.clear();
  }
  return all;
}

// Launch context of a merged run: one stream for every member, and scratch
// for the merged descriptor uploads (reset after each round: the next
// round's uploads are ordered after this round's launches on the stream).
struct MergeCtx {
  hipStream_t s = nullptr;
  Arena sc{size_t(16) << 20};
  static void round_end(void* c) { ((MergeCtx*)c)->sc.reset(); }
};
template <class Batch>
int run_merged(void* const* bs, int n, void* ctx) {
  MergeCtx* c = (MergeCtx*)ctx;
  return Batch::run_many(bs, n, c->sc, c->s);
}
// Run a batch through the calling thread's group, or launch it alone.
template <class Batch>
inline void merge_or_run(Batch* b, Arena& scratch, hipStream_t s) {
  if (MergeGroup* g = merge_group()) {
    g->submit(&run_merged<Batch>, b);
    return;
  }
  void* one = b;
  Batch::run_many(&one, 1, scratch, s);
}

// Every batch type: add() collects descriptors, run() launches them (merged
// with the other members' batches of the same type inside a MergeGroup; an
// empty batch then still takes part in the round, which keeps problems of
// equal structure in step), run_many() is the launch over several batches.

// Every descriptor of one launch must write a distinct output block: tiles of
// different descriptors run concurrently.
template <class T>
struct GemmBatch {
  std::vector<GemmDesc<T>> d;
  // C = alpha op(A) op(B) + beta C ; dims from the blocks
  void add(const Blk<T>& A, int opA, const Blk<T>& B, int opB, const Blk<T>& C, double2 alpha, double2 beta) {
    int m = opA ? A.c : A.r;
    int k = opA ? A.r : A.c;
    int n = opB ? B.r : B.c;
    if (m == 0 || n == 0) return;
    GemmDesc<T> g;
    g.A = A.p;
    g.B = B.p;
    g.C = C.p;
    g.m = m;
    g.n = n;
    g.k = k;
    g.lda = A.ld;
    g.ldb = B.ld;
    g.ldc = C.ld;
    g.opA = opA;
    g.opB = opB;
    g.alpha = alpha;
    g.beta = beta;
    if (k == 0) {  // pure scaling of C
      if (beta.x == 1.0 && beta.y == 0.0) return;
      g.k = 0;
    }
    d.push_back(g);
  }
  void run(Arena& scratch, hipStream_t s) {
    if (d.empty() && !merge_group()) return;
    merge_or_run(this, scratch, s);
  }
  static int run_many(void* const* bs, int n, Arena& scratch, hipStream_t s) {
    std::vector<GemmDesc<T>> d = take_all(bs, n, &GemmBatch::d);
    if (d.empty()) return 0;
    size_t nt = 0;
    for (auto& g : d) nt += (size_t)((g.m + 31) / 32) * ((g.n + 31) / 32);
    std::vector<TileTask> tiles;
    tiles.reserve(nt);
    for (int i = 0; i < (int)d.size(); ++i) {
      int tm = (d[i].m + 31) / 32, tn = (d[i].n + 31) / 32;
      for (int a = 0; a < tm; ++a)
        for (int b = 0; b < tn; ++b) tiles.push_back(TileTask{i, a, b});
    }
    // descriptors and tiles in one copy
    const size_t bd = d.size() * sizeof(GemmDesc<T>), od = (bd + 255) & ~size_t(255);
    char* dev = (char*)scratch.alloc(od + tiles.size() * sizeof(TileTask));
    h2d2(dev, d.data(), bd, tiles.data(), tiles.size() * sizeof(TileTask), s);
    auto* dd = (GemmDesc<T>*)dev;
    auto* dt = (TileTask*)(dev + od);
    double by = 0, fl = 0;
    if (prof_on())
      for (auto& g : d) {
        bool bz = g.beta.x == 0.0 && g.beta.y == 0.0;
        by += sizeof(T) * ((double)g.m * g.k + (double)g.k * g.n + (double)g.m * g.n * (bz ? 1 : 2));
        fl += 8.0 * g.m * g.n * g.k;
      }
    ProfScope ps(s, KNAME(T, "k_gemm"), by, fl);
    launch_gemm<T>((int)tiles.size(), dt, dd, s);
    return 1;
  }
};

template <class T>
struct CopyBatch {
  std::vector<CopyDesc<T>> d;
  void add(const Blk<T>& src, const Blk<T>& dst, int trans = 0, double2 alpha = make_double2(1, 0), int acc = 0) {
    int nr = dst.r, nc = dst.c;
    if (nr == 0 || nc == 0) return;
    CopyDesc<T> c;
    c.src = src.p;
    c.dst = dst.p;
    c.nr = nr;
    c.nc = nc;
    c.lds = src.ld;
    c.ldd = dst.ld;
    c.trans = trans;
    c.accumulate = acc;
    c.alpha = alpha;
    d.push_back(c);
  }
  void run(Arena& scratch, hipStream_t s) {
    if (d.empty() && !merge_group()) return;
    merge_or_run(this, scratch, s);
  }
  static int run_many(void* const* bs, int n, Arena& scratch, hipStream_t s) {
    std::vector<CopyDesc<T>> d = take_all(bs, n, &CopyBatch::d);
    if (d.empty()) return 0;
    auto* dd = upload(scratch, d, s);
    double by = 0;
    if (prof_on())
      for (auto& c : d) by += sizeof(T) * (double)c.nr * c.nc * ((c.trans >= 3 ? 1 : 2) + c.accumulate);
    ProfScope ps(s, KNAME(T, "k_copy"), by, 0);
    launch_copy<T>((int)d.size(), dd, s);
    return 1;
  }
};

template <class T>
struct LUBatch {
  std::vector<LUDesc<T>> d;
  void add(const Blk<T>& A, int* piv, const int* gate = nullptr, int want = 0) {
    if (A.r == 0) return;
    d.push_back(LUDesc<T>{A.p, piv, A.r, A.ld, gate, want});
  }
  void run(Arena& scratch, hipStream_t s) {
    if (d.empty() && !merge_group()) return;
    merge_or_run(this, scratch, s);
  }
  static int run_many(void* const* bs, int n, Arena& scratch, hipStream_t s) {
    std::vector<LUDesc<T>> d = take_all(bs, n, &LUBatch::d);
    if (d.empty()) return 0;
    auto* dd = upload(scratch, d, s);
    double by = 0, fl = 0;
    if (prof_on())
      for (auto& u : d) by += sizeof(T) * 2.0 * u.n * u.n, fl += 8.0 * u.n * u.n * u.n / 3;
    ProfScope ps(s, KNAME(T, "k_lu"), by, fl);
    // blocks that fit the LDS by the LDS kernel, the rest by the global one
    const int lim = lu_lds_max(sizeof(T));
    std::vector<LUDesc<T>> big;
    int nsmall = 0;
    for (auto& u : d) {
      if (u.n > lim) big.push_back(u);
      else nsmall = std::max(nsmall, u.n);
    }
    if (nsmall) launch_lu_lds<T>((int)d.size(), dd, nsmall, s);
    if (!big.empty()) launch_lu<T>((int)big.size(), upload(scratch, big, s), s);
    return (nsmall ? 1 : 0) + (big.empty() ? 0 : 1);
  }
};

template <class T>
struct SolveBatch {
  std::vector<SolveDesc<T>> d;
  int maxc = 0;
  void add(const Blk<T>& LU, const int* piv, const Blk<T>& B, const int* gate = nullptr, int want = 0) {
    if (LU.r == 0 || B.c == 0) return;
    d.push_back(SolveDesc<T>{LU.p, piv, B.p, LU.r, B.c, LU.ld, B.ld, gate, want});
    int ch = (B.c + 63) / 64;
    if (ch > maxc) maxc = ch;
  }
  void run(Arena& scratch, hipStream_t s) {
    maxc = 0;
    if (d.empty() && !merge_group()) return;
    merge_or_run(this, scratch, s);
  }
  static int run_many(void* const* bs, int n, Arena& scratch, hipStream_t s) {
    std::vector<SolveDesc<T>> d = take_all(bs, n, &SolveBatch::d);
    if (d.empty()) return 0;
    auto* dd = upload(scratch, d, s);
    double by = 0, fl = 0;
    if (prof_on())
      for (auto& u : d) by += sizeof(T) * ((double)u.n * u.n + 2.0 * u.n * u.nrhs), fl += 8.0 * u.n * u.n * u.nrhs;
    ProfScope ps(s, KNAME(T, "k_lu_solve"), by, fl);
    const int lim = lu_lds_max(sizeof(T));
    std::vector<SolveDesc<T>> big;
    int nsmall = 0, csmall = 0, cbig = 0;
    for (auto& u : d) {
      if (u.n > lim) {
        big.push_back(u);
        cbig = std::max(cbig, (u.nrhs + 63) / 64);
      } else {
        nsmall = std::max(nsmall, u.n);
        csmall = std::max(csmall, u.nrhs);
      }
    }
    if (nsmall) launch_lu_solve_lds<T>((int)d.size(), csmall, dd, nsmall, s);
    if (!big.empty()) launch_lu_solve<T>((int)big.size(), cbig, upload(scratch, big, s), s);
    return (nsmall ? 1 : 0) + (big.empty() ? 0 : 1);
  }
};

// A square matrix factored in place by column-pivoted QR: the Householder
// vectors and R overwrite the block; tau, perm and the nonzero-pivot count live
// beside it.
template <class T>
struct QRFact {
  Blk<T> A;
  T* tau = nullptr;
  int* perm = nullptr;
  int* nz = nullptr;
  int* rank = nullptr;
};

template <class T>
QRFact<T> new_qr(Arena& a, const Blk<T>& A) {
  QRFact<T> f;
  f.A = A;
  int n = A.c > 0 ? A.c : 1;
  f.tau = a.get<T>(n);
  f.perm = a.get<int>(n);
  f.nz = a.get<int>(2);
  f.rank = f.nz + 1;
  return f;
}

template <class T>
struct QRBatch {
  std::vector<QRDesc<T>> d;
  void add(const QRFact<T>& f, Arena& scratch, double nz_eps = 0) {
    if (f.A.r == 0 || f.A.c == 0) return;
    QRDesc<T> q{};
    q.nz_eps = nz_eps;
    q.A = f.A.p;
    q.tau = f.tau;
    q.perm = f.perm;
    q.work = scratch.get<typename CplxOf<T>::real>(2 * (size_t)f.A.c);
    q.m = f.A.r;
    q.n = f.A.c;
    q.lda = f.A.ld;
    q.thr = 0;
    q.rank_out = f.rank;
    q.maxpiv_out = nullptr;
    q.nz_out = f.nz;
    d.push_back(q);
  }
  void run(Arena& scratch, hipStream_t s) {
    if (d.empty() && !merge_group()) return;
    merge_or_run(this, scratch, s);
  }
  // the matrices that fit k_cpqr_reg first, the rest after; one launch per
  // kernel (AIFMM_CPQR_REG=0: every matrix through k_cpqr)
  static int run_many(void* const* bs, int n, Arena& scratch, hipStream_t s) {
    std::vector<QRDesc<T>> d = take_all(bs, n, &QRBatch::d);
    if (d.empty()) return 0;
    static const bool reg = !getenv("AIFMM_CPQR_REG") || atoi(getenv("AIFMM_CPQR_REG")) != 0;
    const auto small = [](const QRDesc<T>& q) { return cpqr_shape(q.m, q.n) == CPQR_REG; };
    const int ns = reg ? (int)(std::stable_partition(d.begin(), d.end(), small) - d.begin()) : 0;
    auto* dd = upload(scratch, d, s);
    double fl = 0;
    if (prof_on())
      for (auto& q : d) fl += 8.0 * ((double)q.m * q.n * q.n - (double)q.n * q.n * q.n / 3);
    ProfScope ps(s, KNAME(T, "k_cpqr"), 0, fl);
    launch_cpqr<T>(ns, dd, s, CPQR_REG);
    launch_cpqr<T>((int)d.size() - ns, dd + ns, s, CPQR_GLOBAL);
    return 1;
  }
};

// B <- A^{-1} B with a QRFact; B is overwritten (through a scratch copy).
template <class T>
struct QRSolveBatch {
  std::vector<QSolveDesc<T>> d;
  int maxc = 0;
  void add(const QRFact<T>& f, const Blk<T>& B, Arena& scratch) {
    if (f.A.r == 0 || B.c == 0) return;
    QSolveDesc<T> q{};
    q.A = f.A.p;
    q.tau = f.tau;
    q.perm = f.perm;
    q.nz = f.nz;
    q.B = B.p;
    q.scratch = scratch.get<T>((size_t)f.A.r * B.c);
    q.n = f.A.r;
    q.nrhs = B.c;
    q.lda = f.A.ld;
    q.ldb = B.ld;
    d.push_back(q);
    int ch = (B.c + 15) / 16;
    if (ch > maxc) maxc = ch;
  }
  void run(Arena& scratch, hipStream_t s) {
    maxc = 0;
    if (d.empty() && !merge_group()) return;
    merge_or_run(this, scratch, s);
  }
  static int run_many(void* const* bs, int n, Arena& scratch, hipStream_t s) {
    std::vector<QSolveDesc<T>> d = take_all(bs, n, &QRSolveBatch::d);
    if (d.empty()) return 0;
    int maxc = 0;
    for (auto& q : d) maxc = std::max(maxc, (q.nrhs + 15) / 16);
    auto* dd = upload(scratch, d, s);
    double by = 0, fl = 0;
    if (prof_on())
      for (auto& q : d) by += sizeof(T) * ((double)q.n * q.n + 2.0 * q.n * q.nrhs), fl += 16.0 * q.n * q.n * q.nrhs;
    ProfScope ps(s, KNAME(T, "k_qr_solve"), by, fl);
    launch_qr_solve<T>((int)d.size(), maxc, dd, s);
    return 1;
  }
};

// Householder vectors of a column-pivoted QR expanded into the first k
// columns of Q.
template <class T>
struct QFormBatch {
  std::vector<QFormDesc<T>> d;
  void run(Arena& scratch, hipStream_t s) {
    if (d.empty() && !merge_group()) return;
    merge_or_run(this, scratch, s);
  }
  static int run_many(void* const* bs, int n, Arena& scratch, hipStream_t s) {
    std::vector<QFormDesc<T>> d = take_all(bs, n, &QFormBatch::d);
    if (d.empty()) return 0;
    auto* dd = upload(scratch, d, s);
    ProfScope ps(s, KNAME(T, "k_qform"), 0, 0);
    launch_qform<T>((int)d.size(), dd, s);
    return 1;
  }
};

// Frobenius norms: block (src, nr x nc, lds) -> *out.
template <class T>
struct FroBatch {
  std::vector<CopyDesc<T>> d;
  void add(const T* src, int nr, int nc, int lds, double* out) {
    CopyDesc<T> c{};
    c.src = src;
    c.dst = (T*)out;  // the kernel writes the norm through dst
    c.nr = nr;
    c.nc = nc;
    c.lds = lds;
    d.push_back(c);
  }
  void run(Arena& scratch, hipStream_t s) {
    if (d.empty() && !merge_group()) return;
    merge_or_run(this, scratch, s);
  }
  static int run_many(void* const* bs, int n, Arena& scratch, hipStream_t s) {
    std::vector<CopyDesc<T>> d = take_all(bs, n, &FroBatch::d);
    if (d.empty()) return 0;
    auto* dd = upload(scratch, d, s);
    ProfScope ps(s, KNAME(T, "k_fro"), 0, 0);
    launch_fro<T>((int)d.size(), dd, nullptr, s);
    return 1;
  }
};

// Element-wise FP64 -> working precision conversions.
template <class T>
struct CvtBatch {
  std::vector<CvtDesc<T>> d;
  void add(const double2* src, T* dst, int64_t n, const int* gate = nullptr, int want = 0) {
    if (n > 0) d.push_back(CvtDesc<T>{src, dst, n, gate, want});
  }
  void run(Arena& scratch, hipStream_t s) {
    if (d.empty() && !merge_group()) return;
    merge_or_run(this, scratch, s);
  }
  static int run_many(void* const* bs, int n, Arena& scratch, hipStream_t s) {
    std::vector<CvtDesc<T>> d = take_all(bs, n, &CvtBatch::d);
    if (d.empty()) return 0;
    launch_cvt<T>((int)d.size(), upload(scratch, d, s), s, max_len(d));
    return 1;
  }
};

// Element-wise complex64 -> FP64 conversions.
struct WidenBatch {
  std::vector<WidenDesc> d;
  void add(const float2* src, double2* dst, int64_t n, const int* gate = nullptr, int want = 0) {
    if (n > 0) d.push_back(WidenDesc{src, dst, n, gate, want});
  }
  void run(Arena& scratch, hipStream_t s) {
    if (d.empty() && !merge_group()) return;
    merge_or_run(this, scratch, s);
  }
  static int run_many(void* const* bs, int n, Arena& scratch, hipStream_t s) {
    std::vector<WidenDesc> d = take_all(bs, n, &WidenBatch::d);
    if (d.empty()) return 0;
    launch_widen((int)d.size(), upload(scratch, d, s), s, max_len(d));
    return 1;
  }
};

// Complex64 pivot-block LUs whose precision is chosen per block on the
// device (GateDesc).
struct GateBatch {
  std::vector<GateDesc> d;
  static int run_many(void* const* bs, int n, Arena& scratch, hipStream_t s) {
    std::vector<GateDesc> d = take_all(bs, n, &GateBatch::d);
    if (d.empty()) return 0;
    const int lim = lu_lds_max(sizeof(float2));
    int nmax = 0;
    for (auto& g : d)
      if (g.n <= lim) nmax = std::max(nmax, g.n);
    auto* dd = upload(scratch, d, s);
    double by = 0, fl = 0;
    if (prof_on())
      for (auto& g : d) by += 8.0 * 2.0 * g.n * g.n, fl += 8.0 * g.n * g.n * (double)g.n / 3;
    ProfScope ps(s, "k_lu_gate<c64>", by, fl);
    launch_lu_gate((int)d.size(), dd, nmax, s);
    return 1;
  }
  void run(Arena& scratch, hipStream_t s) {
    if (d.empty() && !merge_group()) return;
    merge_or_run(this, scratch, s);
  }
};

// Solves with gated complex64 pivot blocks (MixSolveDesc).
struct MixSolveBatch {
  std::vector<MixSolveDesc> d;
  static int run_many(void* const* bs, int n, Arena& scratch, hipStream_t s) {
    std::vector<MixSolveDesc> d = take_all(bs, n, &MixSolveBatch::d);
    if (d.empty()) return 0;
    int nmax = 0, cmax = 0;
    for (auto& m : d) nmax = std::max(nmax, m.n), cmax = std::max(cmax, m.nrhs);
    auto* dd = upload(scratch, d, s);
    double by = 0, fl = 0;
    if (prof_on())
      for (auto& m : d) by += 8.0 * ((double)m.n * m.n + 2.0 * m.n * m.nrhs), fl += 8.0 * m.n * m.n * (double)m.nrhs;
    ProfScope ps(s, "k_lu_solve_mixed", by, fl);
    launch_lu_solve_mixed((int)d.size(), cmax, dd, nmax, s);
    return 1;
  }
  void run(Arena& scratch, hipStream_t s) {
    if (d.empty() && !merge_group()) return;
    merge_or_run(this, scratch, s);
  }
};

}  // namespace aifmm
