// Box programs: the solve sweeps and the hierarchical matvec expressed as a
// fixed sequence of launches over device-resident descriptors.
//
// One PProg is run by one workgroup.  Phase 1 forms a vector
//     y = src_alpha * src + sum_t alpha_t A_t x_t          (m rows)
// in LDS (optionally storing it to `save`); phase 2 forms up to two outputs
//     out_o = sum_u alpha_u A_u x_u                        (rows_o rows)
// where a term with x == nullptr multiplies the phase-1 vector.  All vectors
// are column-major with nrhs columns.  Every descriptor of one stage writes
// its own output blocks, and reads only blocks no descriptor of the same stage
// writes, so the workgroups of a stage are independent.
#pragma once
#include <vector>

#include "arena.h"
#include "common.h"

namespace aifmm {

template <class T>
struct PTerm {
  const T* A;  // rows x cols, leading dimension lda
  const T* x;  // cols x nrhs, leading dimension ldx; nullptr: the phase-1 vector
  int lda, ldx, cols;
  double2 alpha;
};

template <class T>
struct PProg {
  int m;  // phase-1 rows
  const T* src;
  int ldsrc;
  double2 src_alpha;
  int t0, t1;  // phase-1 terms
  T* save;
  int ldsave;
  int nout;
  T* out[2];
  int ldo[2], rows[2], u0[2], u1[2];
};

// Launch one stage: nprog workgroups; ldy >= every program's m.
template <class T>
void launch_prog(int nprog, const PProg<T>* progs, const PTerm<T>* terms, int nrhs, int ldy, hipStream_t s);
// largest phase-1 length a stage may use
int prog_max_m(int elem_bytes);

template <class T>
class Program {
 public:
  // Start a new stage (a launch boundary).
  void stage() { st_.push_back(Stage{}); }
  // Append a program to the current stage; fill its phase-1 terms with term()
  // between begin_phase1() and end_phase1(), and each output's between
  // begin_out(o) and end_out(o).
  PProg<T>& add(int m) {
    PProg<T> p{};
    p.m = m;
    p.src_alpha = make_double2(1, 0);
    st_.back().progs.push_back(p);
    if (m > st_.back().maxm) st_.back().maxm = m;
    return st_.back().progs.back();
  }
  PProg<T>& cur() { return st_.back().progs.back(); }
  void term(const T* A, int lda, int cols, const T* x, int ldx, double2 alpha) {
    if (cols <= 0) return;
    terms_.push_back(PTerm<T>{A, x, lda, ldx, cols, alpha});
  }
  void begin_phase1() { cur().t0 = (int)terms_.size(); }
  void end_phase1() { cur().t1 = (int)terms_.size(); }
  void begin_out(int o, T* out, int ldo, int rows) {
    PProg<T>& p = cur();
    p.out[o] = out;
    p.ldo[o] = ldo;
    p.rows[o] = rows;
    p.u0[o] = (int)terms_.size();
    if (o + 1 > p.nout) p.nout = o + 1;
  }
  void end_out(int o) { cur().u1[o] = (int)terms_.size(); }

  // Copy descriptors to the device (arena-owned memory).
  void finalize(Arena& a, hipStream_t s);
  // Launch every stage in order on s (asynchronous).
  void run(int nrhs, hipStream_t s) const;
  size_t stages() const { return st_.size(); }
  size_t nterms() const { return terms_.size(); }
  // algorithmic bytes of one run at nrhs (matrices once, vectors per column)
  double bytes(int nrhs) const;
  double flops(int nrhs) const;

 private:
  struct Stage {
    std::vector<PProg<T>> progs;
    int maxm = 0;
    PProg<T>* dprogs = nullptr;
  };
  std::vector<Stage> st_;
  std::vector<PTerm<T>> terms_;
  PTerm<T>* dterms_ = nullptr;
};

}  // namespace aifmm
