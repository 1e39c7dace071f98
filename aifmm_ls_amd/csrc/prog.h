// Box programs: the solve sweeps and the hierarchical matvec expressed as a
// fixed sequence of launches over device-resident descriptors.
//
// One PProg is run by one workgroup.  Phase 1 forms a vector
//     y = src_alpha * src + sum_t alpha_t A_t x_t          (m rows)
// in LDS (optionally storing it to `save`); phase 2 forms up to two outputs
//     out_o = sum_u alpha_u A_u x_u                        (rows_o rows)
// where a term with x == nullptr multiplies the phase-1 vector.  All vectors
// are column-major with nrhs columns.  The matrices A are of the program's
// element type T (complex128 or complex64); the vectors are complex128 and
// every product is formed and summed in FP64, so a complex64 program applies
// its complex64 matrices without rounding its vectors to complex64.  Every descriptor of one stage writes
// its own output blocks, and reads only blocks no descriptor of the same stage
// writes, so the workgroups of a stage are independent.
//
// The builder takes whole box operations (BoxOp) and lowers them: a box with
// few rows becomes one program; a tall box is cut into row tiles run by many
// workgroups over two launches (phase 1 into a vector in memory, then the
// outputs from it), because one workgroup alone cannot keep enough loads in
// flight for it.
#pragma once
#include <atomic>
#include <functional>
#include <map>
#include <cstdint>
#include <vector>

#include "arena.h"
#include "common.h"

namespace aifmm {

template <class T>
struct PTerm {
  const T* A;        // rows x cols, leading dimension lda
  const double2* x;  // cols x nrhs, leading dimension ldx; nullptr: the phase-1 vector
  int lda, ldx, cols;
  double2 alpha;
};

template <class T>
struct PProg {
  int m;  // phase-1 rows
  const double2* src;
  int ldsrc;
  double2 src_alpha;
  int t0, t1;  // phase-1 terms
  double2* save;
  int ldsave;
  int nout;
  double2* out[2];
  int ldo[2], rows[2], u0[2], u1[2];
  int ext;  // in a PSlot: 1 if the term indices refer to the stage-wide array
};

// A program as the kernel reads it: the descriptor and (up to PSLOT_TERMS) its
// terms in one fixed-size slot, so one workgroup fetches all of it in a
// single round of loads (a program with more terms keeps its indices into the
// shared term array, ext = 1).
constexpr int PSLOT_TERMS = 18;
template <class T>
struct PSlot {
  PProg<T> p;
  PTerm<T> t[PSLOT_TERMS];
};

// Launch one stage: nprog workgroups; ldy >= every program's m; the first
// slot_words 16-byte words of each slot hold the descriptor and its terms.
// shape: workgroup threads (256, 512 or 1024; 0 = prog_default_nth(nprog))
// plus PROG_SHAPE_CH x right-hand-side columns per pass (1, 2, 4, or 8 at up
// to 512 threads; 0 = the most of 1, 2, 4 that fit the LDS).
constexpr int PROG_SHAPE_CH = 4096;
// xcols: the most input-vector columns (terms with x, summed) any one sum of
// the stage reads; where they fit the LDS the kernel stages them there.
template <class T>
void launch_prog(int nprog, const PSlot<T>* progs, const PTerm<T>* terms, int nrhs, int ldy, int slot_words,
                 hipStream_t s, int shape = 0, int xcols = 0);
int prog_default_nth(int nprog);
// Launch tuning for programs not yet tuned at an nrhs: 0 off, 1 timed (the
// default unless AIFMM_TUNE=0 is set in the environment), or a launch shape
// (as for launch_prog) to give every stage that shape wherever it fits.
void prog_set_tuning(int mode);
// whether a stage of phase-1 length ldy fits the LDS in this shape
template <class T>
bool prog_shape_fits(int nrhs, int ldy, int shape);
// largest phase-1 length a stage may use (complex128 vectors in LDS)
int prog_max_m();
// while held (counted), programs launch directly instead of through graphs
void prog_graphs_hold(bool on);
// Threads inside the library (one ApiScope per C ABI call, reentrant per
// thread).  A sweep is captured into a HIP graph only while its thread is the
// only one inside, and threads that enter during a capture wait for it to end:
// another thread's frees, synchronisations or copies would otherwise meet the
// capture.
struct ApiScope {
  ApiScope();
  ~ApiScope();
  ApiScope(const ApiScope&) = delete;
  ApiScope& operator=(const ApiScope&) = delete;
};

// One box operation before lowering.
template <class T>
struct BoxOp {
  struct Out {
    double2* out;
    int ldo, rows;
    std::vector<PTerm<T>> terms;
  };
  int m = 0;
  const double2* src = nullptr;
  int ldsrc = 0;
  double2 src_alpha = {1, 0};
  std::vector<PTerm<T>> terms;  // phase 1
  double2* save = nullptr;      // phase-1 result (optional)
  int ldsave = 0;
  double2* ybuf = nullptr;      // where a tiled box keeps y when save is null
  int ldybuf = 0;
  std::vector<Out> outs;        // phase 2 (at most 2)
  bool whole = false;           // one program even when tall (a few boxes whose outputs read all of y)
  int tile = 0;                 // rows per program of a box without outputs (0: the program's default)
  void term(const T* A, int lda, int cols, const double2* x, int ldx, double2 alpha) {
    if (cols > 0) terms.push_back(PTerm<T>{A, x, lda, ldx, cols, alpha});
  }
  Out& out(double2* o, int ldo, int rows) {
    outs.push_back(Out{o, ldo, rows, {}});
    return outs.back();
  }
};

template <class T>
inline void out_term(typename BoxOp<T>::Out& o, const T* A, int lda, int cols, const double2* x, int ldx,
                     double2 alpha) {
  if (cols > 0) o.terms.push_back(PTerm<T>{A, x, lda, ldx, cols, alpha});
}

template <class T>
class Program {
 public:
  // Boxes taller than tile_m() rows are cut into TILE-row tiles over two
  // launches (phase 1 per tile, then the outputs per tile).  The height
  // depends on the right-hand-side capacity the program is built for.  One
  // column: 16 rows, so even the 36-row leaves spread over many short
  // workgroups (measured on MI355X at N = 147,456, complex64 sweep: 64 ->
  // 1.57 ms, 32 -> 1.40 ms, 16 -> 1.35 ms, 128 -> 1.85 ms; tiles of 8, 12,
  // 18 or 32 rows were slower than 16).  Several columns: 64 rows, since
  // each tile's outputs read the box's whole phase-1 vector from memory
  // (16 RHS: 3.26 ms at 64 rows against 3.63 ms at 16; 64 RHS: 8.1 against
  // 10.7 ms).
  static constexpr int TILE = 16;
  static int tile_m(int rhs_capacity) { return rhs_capacity <= 1 ? 16 : 64; }
  explicit Program(int rhs_capacity = 1) : id_(next_id()), tile_m_(tile_m(rhs_capacity)) {}

  // Start a new stage (a launch boundary; a stage with tiled boxes becomes
  // two launches).
  void stage() { close(); }
  // Add one box operation to the current stage.  A tiled box without `save`
  // needs `ybuf` (m x nrhs).
  void add(const BoxOp<T>& op);

  // Copy descriptors to the device (arena-owned memory).
  void finalize(Arena& a, hipStream_t s);
  // Build (before finalize) a program whose stage i runs stage i of every
  // program in ps: the programs belong to independent problems, so their
  // stages can share launches (one level-synchronous launch sequence for a
  // batch of problems).
  void merge_from(const std::vector<const Program<T>*>& ps);
  // Launch every stage in order on s (asynchronous).  The launch sequence is
  // captured once per nrhs into a HIP graph and replayed; with kernel
  // profiling on, the launches are issued one by one instead.
  //
  // With `reload` (which re-issues the copy of the program's input on s), the
  // first run at each nrhs first tunes the launch shapes: see tune().
  void run(int nrhs, hipStream_t s, const std::function<void()>& reload = nullptr);
  // One run of a program that runs once: the launches issued one by one (no
  // graph to capture and instantiate), no tuning.
  void run_once(int nrhs, hipStream_t s) const { launch_all(nrhs, s, false); }
  // Outcome of tuning at one nrhs: mean milliseconds of a whole run with the
  // default and with the picked workgroup sizes, stages given another size,
  // and whether the picked sizes were kept.
  struct TuneResult {
    double default_ms = 0, tuned_ms = 0;
    int changed = 0;
    bool adopted = false;
  };
  // Give every stage `shape` (where it fits) at nrhs instead of tuning (test
  // hook; a graph captured at nrhs is dropped).
  void force_shape(int nrhs, int shape);
  const TuneResult* tune_result(int nrhs) const {
    auto it = tune_.find(nrhs);
    return it == tune_.end() ? nullptr : &it->second;
  }
  // the most recent tuning (any nrhs), or null
  const TuneResult* last_tune() const { return last_tune_; }

  // distinct for every program built in the process (cache keys)
  uint64_t id() const { return id_; }
  Program(const Program&) = delete;
  Program& operator=(const Program&) = delete;
  ~Program();
  size_t stages() const { return st_.size(); }
  size_t nterms() const { return terms_.size(); }
  // algorithmic bytes of one run at nrhs (matrices once, vectors per column)
  double bytes(int nrhs) const;
  double flops(int nrhs) const;

 private:
  struct Stage {
    std::vector<PProg<T>> progs;
    int maxm = 0;
    PSlot<T>* dslots = nullptr;
    int slot_words = 0;
    int xcols = 0;  // input-vector columns of the stage's largest sum
  };
  static uint64_t next_id() {
    static std::atomic<uint64_t> n{0};
    return ++n;
  }
  const uint64_t id_;
  const int tile_m_;
  std::vector<Stage> st_;
  std::vector<PTerm<T>> terms_;
  PTerm<T>* dterms_ = nullptr;
  hipGraphExec_t gexec_ = nullptr;
  int gnrhs_ = -1;
  // workgroup size per stage and nrhs (0 = the default size)
  std::map<int, std::vector<int>> shapes_;
  std::map<int, TuneResult> tune_;
  const TuneResult* last_tune_ = nullptr;
  void launch_all(int nrhs, hipStream_t s, bool prof) const;
  // Launch autotuner: the whole program is run (on freshly reloaded input)
  // once per candidate launch shape (256, 512, 1024 threads; with several
  // right-hand sides also 1, 2, 4 or 8 columns per pass; and the default)
  // with an event after every stage; each stage takes the shape that
  // measured fastest (by at least 3 % over the default).  The picked shapes
  // are kept only if whole runs with them beat the default by 1 %.  Whole
  // runs rather than stages in isolation: a stage repeated alone finds its
  // data in cache and times unlike its place in the sequence.
  void tune(int nrhs, hipStream_t s, const std::function<void()>& reload);
  // pending launches of the open stage: A = fused programs and tile phase 1,
  // B = tile phase 2
  Stage pa_, pb_;
  void push(Stage& g, const PProg<T>& p) {
    g.progs.push_back(p);
    if (p.m > g.maxm) g.maxm = p.m;
  }
  void close() {
    if (!pa_.progs.empty()) st_.push_back(pa_);
    if (!pb_.progs.empty()) st_.push_back(pb_);
    pa_ = Stage{};
    pb_ = Stage{};
  }
};

}  // namespace aifmm
