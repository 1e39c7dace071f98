// Reader and writer for the reference's rhs_x/*.bin vectors: a cereal
// PortableBinary archive of one std::vector<double> (testFMM2D_find_error.cpp:
// 19-37, storedata::load_vec / save_vec).  Layout: one byte holding the
// writer's endianness (1 = little), the element count as uint64, then the
// doubles in the writer's byte order.
#pragma once
#include <cstdint>
#include <cstring>
#include <fstream>
#include <stdexcept>
#include <string>
#include <vector>

namespace aifmm {

inline bool host_little_endian() {
  const uint16_t one = 1;
  unsigned char b;
  std::memcpy(&b, &one, 1);
  return b == 1;
}

template <class T>
inline T byteswap_value(T v) {
  unsigned char b[sizeof(T)];
  std::memcpy(b, &v, sizeof(T));
  for (size_t i = 0; i < sizeof(T) / 2; ++i) std::swap(b[i], b[sizeof(T) - 1 - i]);
  std::memcpy(&v, b, sizeof(T));
  return v;
}

inline std::vector<double> load_vec(const std::string& fname) {
  std::ifstream in(fname, std::ios::binary);
  if (!in) throw std::runtime_error("cannot open " + fname);
  uint8_t little = 0;
  in.read((char*)&little, 1);
  const bool swap = (little != 0) != host_little_endian();
  uint64_t n = 0;
  in.read((char*)&n, 8);
  if (swap) n = byteswap_value(n);
  std::vector<double> v(n);
  if (n) in.read((char*)v.data(), (std::streamsize)(n * 8));
  if (!in) throw std::runtime_error("truncated vector file " + fname);
  if (swap)
    for (auto& d : v) d = byteswap_value(d);
  return v;
}

inline void save_vec(const std::string& fname, const std::vector<double>& v) {
  std::ofstream out(fname, std::ios::binary);
  if (!out) throw std::runtime_error("cannot write " + fname);
  const uint8_t little = host_little_endian() ? 1 : 0;
  out.write((const char*)&little, 1);
  const uint64_t n = v.size();
  out.write((const char*)&n, 8);
  if (n) out.write((const char*)v.data(), (std::streamsize)(n * 8));
}

}  // namespace aifmm
