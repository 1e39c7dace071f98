// Host side of the box programs: descriptor upload, launch sequence and the
// algorithmic byte / flop count of one run.
#include "prog.h"

#include <condition_variable>
#include <mutex>

#include <algorithm>
#include <cstddef>
#include <atomic>
#include <cstdlib>
#include <string>
#include <type_traits>

#include "batch.h"
#include "prof.h"

namespace aifmm {

template <class T>
void Program<T>::add(const BoxOp<T>& op) {
  if (op.outs.size() > 2) throw std::runtime_error("box program with more than two outputs");
  int maxrows = 0;
  bool uses_y = false;
  for (const auto& o : op.outs) {
    maxrows = std::max(maxrows, o.rows);
    for (const auto& t : o.terms) uses_y |= t.x == nullptr;
  }
  if (op.tile > 0 && !op.outs.empty()) throw std::runtime_error("row tiles of a chosen height need a box without outputs");
  const int tm = op.tile > 0 ? op.tile : tile_m_;
  if ((op.m <= tm && maxrows <= tm) || (op.whole && op.m <= 1024)) {
    PProg<T> p{};
    p.m = op.m;
    p.src = op.src;
    p.ldsrc = op.ldsrc;
    p.src_alpha = op.src_alpha;
    p.save = op.save;
    p.ldsave = op.ldsave;
    p.t0 = (int)terms_.size();
    terms_.insert(terms_.end(), op.terms.begin(), op.terms.end());
    p.t1 = (int)terms_.size();
    p.nout = (int)op.outs.size();
    for (int o = 0; o < p.nout; ++o) {
      p.out[o] = op.outs[o].out;
      p.ldo[o] = op.outs[o].ldo;
      p.rows[o] = op.outs[o].rows;
      p.u0[o] = (int)terms_.size();
      terms_.insert(terms_.end(), op.outs[o].terms.begin(), op.outs[o].terms.end());
      p.u1[o] = (int)terms_.size();
    }
    push(pa_, p);
    return;
  }
  // tiled: phase 1 in row tiles into y (save or ybuf), then the outputs in
  // row tiles reading y from memory
  double2* y = op.save ? op.save : op.ybuf;
  const int ldy = op.save ? op.ldsave : op.ldybuf;
  if (op.m > 0 && !y) throw std::runtime_error("tiled box program without a vector for phase 1");
  if (uses_y && op.m == 0) throw std::runtime_error("box program term reads an empty phase-1 vector");
  const int th = op.tile > 0 ? op.tile : TILE;
  for (int r0 = 0; r0 < op.m; r0 += th) {
    PProg<T> p{};
    p.m = std::min(th, op.m - r0);
    p.src = op.src ? op.src + r0 : nullptr;
    p.ldsrc = op.ldsrc;
    p.src_alpha = op.src_alpha;
    p.save = y + r0;
    p.ldsave = ldy;
    p.t0 = (int)terms_.size();
    for (auto t : op.terms) {
      t.A += r0;
      terms_.push_back(t);
    }
    p.t1 = (int)terms_.size();
    push(pa_, p);
  }
  for (const auto& o : op.outs)
    for (int r0 = 0; r0 < o.rows; r0 += TILE) {
      PProg<T> p{};
      p.m = 0;
      p.nout = 1;
      p.out[0] = o.out + r0;
      p.ldo[0] = o.ldo;
      p.rows[0] = std::min(TILE, o.rows - r0);
      p.u0[0] = (int)terms_.size();
      for (auto t : o.terms) {
        t.A += r0;
        if (!t.x) {
          t.x = y;
          t.ldx = ldy;
        }
        terms_.push_back(t);
      }
      p.u1[0] = (int)terms_.size();
      push(pb_, p);
    }
}

template <class T>
void Program<T>::merge_from(const std::vector<const Program<T>*>& ps) {
  size_t ns = 0;
  for (const Program<T>* p : ps) ns = std::max(ns, p->st_.size());
  st_.assign(ns, Stage{});
  for (const Program<T>* p : ps) {
    const int off = (int)terms_.size();
    terms_.insert(terms_.end(), p->terms_.begin(), p->terms_.end());
    for (size_t i = 0; i < p->st_.size(); ++i) {
      Stage& g = st_[i];
      for (PProg<T> q : p->st_[i].progs) {
        q.t0 += off;
        q.t1 += off;
        for (int o = 0; o < q.nout; ++o) {
          q.u0[o] += off;
          q.u1[o] += off;
        }
        g.progs.push_back(q);
      }
      g.maxm = std::max(g.maxm, p->st_[i].maxm);
    }
  }
}

template <class T>
void Program<T>::finalize(Arena& a, hipStream_t s) {
  close();
  const int lim = prog_max_m();
  if (!terms_.empty()) {
    dterms_ = a.get<PTerm<T>>(terms_.size());
    h2d(dterms_, terms_.data(), terms_.size() * sizeof(PTerm<T>), s);
  }
  for (auto& g : st_) {
    if (g.maxm > lim)
      throw std::runtime_error("box program of " + std::to_string(g.maxm) + " rows exceeds the LDS limit " +
                               std::to_string(lim));
    if (g.progs.empty()) continue;
    g.xcols = 0;
    auto xcols = [&](int t0, int t1) {
      int n = 0;
      for (int t = t0; t < t1; ++t)
        if (terms_[t].x) n += terms_[t].cols;
      return n;
    };
    for (const auto& q : g.progs) {
      g.xcols = std::max(g.xcols, xcols(q.t0, q.t1));
      for (int o = 0; o < q.nout; ++o) g.xcols = std::max(g.xcols, xcols(q.u0[o], q.u1[o]));
    }
    std::vector<PSlot<T>> sl(g.progs.size());
    int inl = 0;  // most terms any program of the stage keeps in its slot
    for (size_t i = 0; i < g.progs.size(); ++i) {
      PProg<T> q = g.progs[i];
      int n = q.t1 - q.t0;
      for (int o = 0; o < q.nout; ++o) n += q.u1[o] - q.u0[o];
      q.ext = n > PSLOT_TERMS;
      if (!q.ext) inl = std::max(inl, n);
      if (!q.ext) {
        // terms copied into the slot, indices rebased to it
        int k = 0;
        auto take = [&](int& lo, int& hi) {
          const int b = k;
          for (int t = lo; t < hi; ++t) sl[i].t[k++] = terms_[t];
          lo = b;
          hi = k;
        };
        take(q.t0, q.t1);
        for (int o = 0; o < q.nout; ++o) take(q.u0[o], q.u1[o]);
      }
      sl[i].p = q;
    }
    g.slot_words = (int)((offsetof(PSlot<T>, t) + inl * sizeof(PTerm<T>) + 15) / 16);
    g.dslots = a.get<PSlot<T>>(sl.size());
    h2d(g.dslots, sl.data(), sl.size() * sizeof(PSlot<T>), s);
  }
  HIP_OK(hipStreamSynchronize(s));
}

template <class T>
static double stage_bytes(const std::vector<PProg<T>>& ps, const std::vector<PTerm<T>>& ts, int nrhs) {
  double b = 0;
  const double vb = sizeof(double2);  // vectors are complex128
  auto term_bytes = [&](int t0, int t1, int rows) {
    for (int t = t0; t < t1; ++t) {
      b += (double)rows * ts[t].cols * sizeof(T);
      if (ts[t].x) b += (double)ts[t].cols * nrhs * vb;
    }
  };
  for (auto& p : ps) {
    term_bytes(p.t0, p.t1, p.m);
    if (p.src) b += (double)p.m * nrhs * vb;
    if (p.save) b += (double)p.m * nrhs * vb;
    for (int o = 0; o < p.nout; ++o) {
      term_bytes(p.u0[o], p.u1[o], p.rows[o]);
      b += (double)p.rows[o] * nrhs * vb;
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
void Program<T>::launch_all(int nrhs, hipStream_t s, bool prof) const {
  auto sh = shapes_.find(nrhs);
  const std::vector<int>* shape = sh == shapes_.end() ? nullptr : &sh->second;
  for (size_t k = 0; k < st_.size(); ++k) {
    const Stage& g = st_[k];
    if (g.progs.empty()) continue;
    ProfScope ps(s, std::is_same<T, float2>::value ? "k_prog<c64>" : "k_prog<c128>",
                 prof ? stage_bytes(g.progs, terms_, nrhs) : 0, prof ? stage_flops(g.progs, terms_, nrhs) : 0);
    launch_prog<T>((int)g.progs.size(), g.dslots, dterms_, nrhs, g.maxm, g.slot_words, s, shape ? (*shape)[k] : 0,
                   g.xcols);
  }
}

// 0 off, 1 timed tuning (the default unless AIFMM_TUNE=0 is set), or a fixed
// launch shape (prog.h) for every stage it fits
static std::atomic<int> tuning_mode{-1};
static int tuning() {
  int m = tuning_mode.load();
  if (m < 0) {
    const char* e = getenv("AIFMM_TUNE");
    int expect = -1;
    tuning_mode.compare_exchange_strong(expect, (e && e[0] == '0') ? 0 : 1);
    m = tuning_mode.load();
  }
  return m;
}
void prog_set_tuning(int mode) {
  const int nth = mode % PROG_SHAPE_CH, ch = mode / PROG_SHAPE_CH;
  const bool shape = (nth == 256 || nth == 512 || nth == 1024) &&
                     (ch == 0 || ch == 1 || ch == 2 || ch == 4 || (ch == 8 && nth <= 512));
  if (mode != 0 && mode != 1 && !shape)
    throw std::runtime_error("tuning mode must be 0, 1 or a launch shape (256, 512 or 1024 threads + 4096 x "
                             "columns per pass, 8 columns at up to 512 threads)");
  tuning_mode = mode;
}

template <class T>
void Program<T>::tune(int nrhs, hipStream_t s, const std::function<void()>& reload) {
  TuneResult& R = tune_[nrhs];
  last_tune_ = &R;
  const int ng = (int)st_.size();
  const std::vector<int> def(ng, 0);
  shapes_[nrhs] = def;
  const int mode = tuning();
  if (!mode || !ng) return;
  if (mode > 1) {
    // a fixed shape wherever it fits (exercises every shape deterministically)
    std::vector<int> shape(ng, 0);
    for (int k = 0; k < ng; ++k)
      if (!st_[k].progs.empty() && prog_shape_fits<T>(nrhs, st_[k].maxm, mode)) {
        shape[k] = mode;
        R.changed += mode % PROG_SHAPE_CH != prog_default_nth((int)st_[k].progs.size()) || mode / PROG_SHAPE_CH;
      }
    shapes_[nrhs] = shape;
    R.adopted = true;
    R.default_ms = R.tuned_ms = 1;
    return;
  }
  std::vector<hipEvent_t> ev(ng + 1, nullptr);
  for (auto& e : ev) HIP_OK(hipEventCreate(&e));
  // one whole run on reloaded input; with `per`, the time of every stage
  auto timed = [&](const std::vector<int>& shape, std::vector<float>* per) {
    reload();
    HIP_OK(hipEventRecord(ev[0], s));
    for (int k = 0; k < ng; ++k) {
      const Stage& g = st_[k];
      if (!g.progs.empty())
        launch_prog<T>((int)g.progs.size(), g.dslots, dterms_, nrhs, g.maxm, g.slot_words, s, shape[k], g.xcols);
      if (per || k == ng - 1) HIP_OK(hipEventRecord(ev[k + 1], s));
    }
    HIP_OK(hipEventSynchronize(ev[ng]));
    float tot = 0;
    HIP_OK(hipEventElapsedTime(&tot, ev[0], ev[ng]));
    if (per) {
      per->assign(ng, 0.f);
      for (int k = 0; k < ng; ++k) HIP_OK(hipEventElapsedTime(&(*per)[k], ev[k], ev[k + 1]));
    }
    return tot;
  };
  try {
    // About a dozen whole runs: worth it for programs that are run many times
    // (a refined solve runs its sweep and operator several times per call).
    // A run longer than 20 ms (many right-hand sides of a large problem)
    // keeps the default sizes, since tuning would cost more than it returns.
    const float t0 = timed(def, nullptr);
    const int reps = t0 > 20.f ? 0 : 2;
    if (reps) {
      std::vector<float> best(ng, 3.0e38f), per;
      std::vector<int> pick(ng, 0);
      // candidates: every workgroup size, and with several right-hand sides
      // every count of columns per pass, after the default (0)
      std::vector<int> cands{0};
      for (int nth : {1024, 512, 256})
        for (int ch : {0, 8, 4, 2, 1}) {
          if (ch && nrhs == 1) continue;
          cands.push_back(nth + PROG_SHAPE_CH * ch);
        }
      for (int c : cands) {
        std::vector<int> shape(ng, 0);
        for (int k = 0; k < ng; ++k)
          if (c && prog_shape_fits<T>(nrhs, st_[k].maxm, c)) shape[k] = c;
        std::vector<float> mn(ng, 3.0e38f);
        for (int r = 0; r < reps; ++r) {
          timed(shape, &per);
          for (int k = 0; k < ng; ++k) mn[k] = std::min(mn[k], per[k]);
        }
        // the default (c = 0) is measured first; another shape must beat it by 3 %
        for (int k = 0; k < ng; ++k)
          if (mn[k] < (c ? 0.97f : 1.f) * best[k]) {
            best[k] = mn[k];
            pick[k] = shape[k];
          }
      }
      double td = 0, tt = 0;
      for (int r = 0; r < 3; ++r) {
        td += timed(def, nullptr);
        tt += timed(pick, nullptr);
      }
      R.default_ms = td / 3;
      R.tuned_ms = tt / 3;
      for (int k = 0; k < ng; ++k)
        R.changed += pick[k] && !st_[k].progs.empty() &&
                     (pick[k] % PROG_SHAPE_CH != prog_default_nth((int)st_[k].progs.size()) ||
                      pick[k] / PROG_SHAPE_CH != 0);
      if (tt < 0.99 * td) {
        shapes_[nrhs] = pick;
        R.adopted = true;
      }
    } else {
      R.default_ms = R.tuned_ms = t0;
    }
  } catch (...) {
    for (auto& e : ev) (void)hipEventDestroy(e);
    throw;
  }
  for (auto& e : ev) HIP_OK(hipEventDestroy(e));
}

// AIFMM_GRAPH=0 issues the launches one by one instead of replaying a graph
// (for tools that do not follow stream capture); so does a serving loop,
// whose threads free memory and synchronise while others would capture
static std::atomic<int> graph_holds{0};
void prog_graphs_hold(bool on) { graph_holds += on ? 1 : -1; }

namespace {
std::mutex gate_mu;
std::condition_variable gate_cv;
int gate_inside = 0;      // threads inside the library
bool gate_capturing = false;
thread_local int gate_depth = 0;
}  // namespace

ApiScope::ApiScope() {
  if (gate_depth++) return;
  std::unique_lock<std::mutex> lk(gate_mu);
  gate_cv.wait(lk, [] { return !gate_capturing; });
  ++gate_inside;
}

ApiScope::~ApiScope() {
  if (--gate_depth) return;
  std::lock_guard<std::mutex> g(gate_mu);
  --gate_inside;
  gate_cv.notify_all();
}

// true: the calling thread is alone inside the library and others are held
// out until capture_end()
static bool capture_begin() {
  std::lock_guard<std::mutex> g(gate_mu);
  if (gate_inside > 1 || gate_capturing) return false;
  gate_capturing = true;
  return true;
}

static void capture_end() {
  std::lock_guard<std::mutex> g(gate_mu);
  gate_capturing = false;
  gate_cv.notify_all();
}
static bool graphs_enabled() {
  static const bool on = [] {
    const char* e = getenv("AIFMM_GRAPH");
    return !(e && e[0] == '0');
  }();
  return on && graph_holds.load() == 0;
}

template <class T>
void Program<T>::force_shape(int nrhs, int shape) {
  std::vector<int> sh(st_.size(), 0);
  for (size_t k = 0; k < st_.size(); ++k)
    if (!st_[k].progs.empty() && shape && prog_shape_fits<T>(nrhs, st_[k].maxm, shape)) sh[k] = shape;
  shapes_[nrhs] = sh;
  TuneResult& R = tune_[nrhs];
  R = TuneResult{1, 1, 0, true};
  last_tune_ = &R;
  if (gnrhs_ == nrhs) gnrhs_ = -1;
}

template <class T>
void Program<T>::run(int nrhs, hipStream_t s, const std::function<void()>& reload) {
  if (reload && !tune_.count(nrhs)) {
    tune(nrhs, s, reload);
    reload();
    // a graph captured earlier at this nrhs holds the default sizes
    if (gnrhs_ == nrhs) gnrhs_ = -1;
  }
  if (prof_on() || !graphs_enabled()) {
    launch_all(nrhs, s, prof_on());
    return;
  }
  if (!gexec_ || gnrhs_ != nrhs) {
    if (!capture_begin()) {  // other threads are inside: launch directly this time
      launch_all(nrhs, s, false);
      return;
    }
    hipGraph_t g = nullptr;
    try {
      if (gexec_) HIP_OK(hipGraphExecDestroy(gexec_));
      gexec_ = nullptr;
      HIP_OK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
      try {
        launch_all(nrhs, s, false);
      } catch (...) {
        (void)hipStreamEndCapture(s, &g);
        throw;
      }
      HIP_OK(hipStreamEndCapture(s, &g));
      HIP_OK(hipGraphInstantiate(&gexec_, g, nullptr, nullptr, 0));
      HIP_OK(hipGraphDestroy(g));
      g = nullptr;
    } catch (...) {
      if (g) (void)hipGraphDestroy(g);
      capture_end();
      throw;
    }
    capture_end();
    gnrhs_ = nrhs;
  }
  HIP_OK(hipGraphLaunch(gexec_, s));
}

template <class T>
Program<T>::~Program() {
  if (gexec_) (void)hipGraphExecDestroy(gexec_);
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
