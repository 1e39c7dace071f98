// The dense product y = A x without device allocation: the variant of
// aifmm_ls_dense_matvec for callers that cannot allocate device memory (code
// under HIP graph capture, tight serving loops, harnesses that forbid it).
// The caller passes device buffers for x and y; the product's scratch (leaf
// moments and per-chunk partial sums) lives in the file-scope device array
// `scratch` below, 64 MiB per GPU for as long as the process runs.  Every call
// on one GPU shares that array: calls that overlap on one GPU race.  The
// kernels write every moment and every partial sum before reading it, so the
// contents a previous call left need no clearing.
#include <mutex>
#include <stdexcept>
#include <string>

#include "../../include/aifmm_ls.h"
#include "kernels.h"
#include "problem.h"
#include "prog.h"

namespace aifmm {
void set_last_error(const std::string& msg);   // capi.cpp
Problem* problem_of(aifmm_ls_problem* h);      // capi.cpp
}  // namespace aifmm

namespace {

constexpr size_t kScratchElems = (size_t(64) << 20) / sizeof(double2);  // 64 MiB
static __device__ __align__(256) double2 scratch[kScratchElems];

// The array's device address on each GPU (hipGetSymbolAddress, once per
// device with that device current); host code uses these, never `scratch`.
constexpr int kMaxDevices = 64;
std::mutex addr_mu;
double2* addr[kMaxDevices] = {};

double2* scratch_on_current_device() {
  int dev = -1;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices)
    throw std::runtime_error("no current HIP device");
  std::lock_guard<std::mutex> g(addr_mu);
  if (!addr[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      throw std::runtime_error(std::string("hipGetSymbolAddress(scratch) failed: ") + hipGetErrorString(e));
    addr[dev] = (double2*)p;
  }
  return addr[dev];
}

}  // namespace

extern "C" int aifmm_ls_dense_matvec_noalloc(aifmm_ls_problem* h, int nrhs, const double* x, double* y,
                                              void* stream) {
  try {
    aifmm::ApiScope api_scope_;
    using namespace aifmm;
    if (!h || !x || !y) throw std::runtime_error("null argument");
    if (nrhs < 1) throw std::runtime_error("nrhs must be >= 1");
    Problem& P = *problem_of(h);
    const int64_t N = P.N;
    const size_t moments = (size_t)NTERM * P.nleaf;
    if (moments + (size_t)N > kScratchElems)
      throw std::runtime_error("dense product without allocation: N = " + std::to_string(N) + " needs " +
                               std::to_string((moments + (size_t)N) * sizeof(double2)) +
                               " scratch bytes, above the 64 MiB static buffer");
    int nch = dmv_chunks(N, P.nleaf);
    const int fit = (int)((kScratchElems - moments) / (size_t)N);
    if (nch > fit) nch = fit;
    double2* sc = scratch_on_current_device();
    hipStream_t s = stream ? (hipStream_t)stream : P.stream;
    // the problem's tables and points were uploaded on its own stream
    if (stream) HIP_OK(hipStreamSynchronize(P.stream));
    for (int c = 0; c < nrhs; ++c)
      launch_dense_matvec(P.ctx_phi(), N, (const double2*)x + (size_t)N * c, (double2*)y + (size_t)N * c, sc, nch,
                          s, false);
    if (!stream) HIP_OK(hipStreamSynchronize(s));
    return 0;
  } catch (const std::exception& e) {
    aifmm::set_last_error(e.what());
    return -1;
  } catch (...) {
    aifmm::set_last_error("unknown error");
    return -2;
  }
}
