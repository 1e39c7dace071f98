// Device memory arena: large hipMalloc chunks handed out with 256-byte
// alignment and released together.  Factorisations allocate every block
// from their own arena, so freeing a factorisation is one call.
#pragma once
#include <utility>
#include <vector>

#include "common.h"

namespace aifmm {

class Arena {
 public:
  explicit Arena(size_t chunk = size_t(1) << 30) : chunk_(chunk) {}
  // chunk size of later allocations (problems size their arenas by N)
  void set_chunk(size_t chunk) { chunk_ = chunk; }
  ~Arena() { release(); }
  Arena(const Arena&) = delete;
  Arena& operator=(const Arena&) = delete;

  void* alloc(size_t bytes) {
    bytes = (bytes + 255) & ~size_t(255);
    if (bytes == 0) bytes = 256;
    if (cur_ < chunks_.size() && used_ + bytes <= sizes_[cur_]) {
      void* p = (char*)chunks_[cur_] + used_;
      used_ += bytes;
      total_ += bytes;
      return p;
    }
    // next chunk (reuse a retained one if big enough)
    while (cur_ + 1 < chunks_.size()) {
      ++cur_;
      used_ = 0;
      if (bytes <= sizes_[cur_]) return alloc(bytes);
    }
    size_t sz = bytes > chunk_ ? bytes : chunk_;
    void* p = nullptr;
    HIP_OK(hipMalloc(&p, sz));
    chunks_.push_back(p);
    sizes_.push_back(sz);
    cur_ = chunks_.size() - 1;
    used_ = 0;
    return alloc(bytes);
  }
  template <class T>
  T* get(size_t n) {
    return (T*)alloc(n * sizeof(T));
  }
  // Keep the chunks, forget the contents.
  void reset() {
    cur_ = 0;
    used_ = 0;
    total_ = 0;
  }
  void release() {
    for (void* p : chunks_) (void)hipFree(p);
    chunks_.clear();
    sizes_.clear();
    cur_ = 0;
    used_ = 0;
    total_ = 0;
  }
  size_t reserved() const {
    size_t s = 0;
    for (size_t z : sizes_) s += z;
    return s;
  }
  size_t used() const { return total_; }
  void swap(Arena& o) {
    std::swap(chunk_, o.chunk_);
    chunks_.swap(o.chunks_);
    sizes_.swap(o.sizes_);
    std::swap(cur_, o.cur_);
    std::swap(used_, o.used_);
    std::swap(total_, o.total_);
  }

  // chunk size for an arena holding about per_unknown bytes per unknown
  static size_t chunk_for(int64_t n, size_t per_unknown, size_t lo, size_t hi) {
    const size_t want = (size_t)n * per_unknown;
    return want < lo ? lo : want > hi ? hi : want;
  }

 private:
  size_t chunk_;
  std::vector<void*> chunks_;
  std::vector<size_t> sizes_;
  size_t cur_ = 0, used_ = 0, total_ = 0;
};

}  // namespace aifmm
