// Host side of the box programs: descriptor upload, launch sequence and the
// algorithmic byte / flop count of one run.
#include "prog.h"

#include <string>
#include <type_traits>

#include "prof.h"

namespace aifmm {

template <class T>
void Program<T>::finalize(Arena& a, hipStream_t s) {
  const int lim = prog_max_m(sizeof(T));
  if (!terms_.empty()) {
    dterms_ = a.get<PTerm<T>>(terms_.size());
    HIP_OK(hipMemcpyAsync(dterms_, terms_.data(), terms_.size() * sizeof(PTerm<T>), hipMemcpyHostToDevice, s));
  }
  for (auto& g : st_) {
    if (g.maxm > lim)
      throw std::runtime_error("box program of " + std::to_string(g.maxm) + " rows exceeds the LDS limit " +
                               std::to_string(lim));
    if (g.progs.empty()) continue;
    g.dprogs = a.get<PProg<T>>(g.progs.size());
    HIP_OK(hipMemcpyAsync(g.dprogs, g.progs.data(), g.progs.size() * sizeof(PProg<T>), hipMemcpyHostToDevice, s));
  }
  HIP_OK(hipStreamSynchronize(s));
}

template <class T>
static double stage_bytes(const std::vector<PProg<T>>& ps, const std::vector<PTerm<T>>& ts, int nrhs) {
  double b = 0;
  auto term_bytes = [&](int t0, int t1, int rows) {
    for (int t = t0; t < t1; ++t) {
      b += (double)rows * ts[t].cols * sizeof(T);
      if (ts[t].x) b += (double)ts[t].cols * nrhs * sizeof(T);
    }
  };
  for (auto& p : ps) {
    term_bytes(p.t0, p.t1, p.m);
    if (p.src) b += (double)p.m * nrhs * sizeof(T);
    if (p.save) b += (double)p.m * nrhs * sizeof(T);
    for (int o = 0; o < p.nout; ++o) {
      term_bytes(p.u0[o], p.u1[o], p.rows[o]);
      b += (double)p.rows[o] * nrhs * sizeof(T);
    }
  }
  return b;
}

template <class T>
static double stage_flops(const std::vector<PProg<T>>& ps, const std::vector<PTerm<T>>& ts, int nrhs) {
  double f = 0;
  for (auto& p : ps) {
    for (int t = p.t0; t < p.t1; ++t) f += 8.0 * p.m * ts[t].cols * nrhs;
    for (int o = 0; o < p.nout; ++o)
      for (int t = p.u0[o]; t < p.u1[o]; ++t) f += 8.0 * p.rows[o] * ts[t].cols * nrhs;
  }
  return f;
}

template <class T>
void Program<T>::run(int nrhs, hipStream_t s) const {
  const bool prof = prof_on();
  for (auto& g : st_) {
    if (g.progs.empty()) continue;
    ProfScope ps(s, std::is_same<T, float2>::value ? "k_prog<c64>" : "k_prog<c128>",
                 prof ? stage_bytes(g.progs, terms_, nrhs) : 0, prof ? stage_flops(g.progs, terms_, nrhs) : 0);
    launch_prog<T>((int)g.progs.size(), g.dprogs, dterms_, nrhs, g.maxm, s);
  }
}

template <class T>
double Program<T>::bytes(int nrhs) const {
  double b = 0;
  for (auto& g : st_) b += stage_bytes(g.progs, terms_, nrhs);
  return b;
}

template <class T>
double Program<T>::flops(int nrhs) const {
  double f = 0;
  for (auto& g : st_) f += stage_flops(g.progs, terms_, nrhs);
  return f;
}

template class Program<double2>;
template class Program<float2>;

}  // namespace aifmm
