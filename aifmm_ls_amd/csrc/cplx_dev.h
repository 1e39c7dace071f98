// Device-side complex helpers shared by the kernel files (double2 = complex128,
// float2 = complex64).
#pragma once
#include "common.h"

namespace aifmm {

template <class T>
__device__ __forceinline__ T czero();
template <>
__device__ __forceinline__ double2 czero<double2>() {
  return mk(0.0, 0.0);
}
template <>
__device__ __forceinline__ float2 czero<float2>() {
  return mkf(0.f, 0.f);
}
template <class T>
__device__ __forceinline__ T from2(double2 a);
template <>
__device__ __forceinline__ double2 from2<double2>(double2 a) {
  return a;
}
template <>
__device__ __forceinline__ float2 from2<float2>(double2 a) {
  return mkf((float)a.x, (float)a.y);
}
// a / b by Smith's algorithm (no |b|^2, so no underflow for small b)
template <class T>
__device__ __forceinline__ T cdiv(T a, T b) {
  using R = typename CplxOf<T>::real;
  T r;
  if (fabs(b.x) >= fabs(b.y)) {
    R q = b.y / b.x, den = b.x + b.y * q;
    r.x = (a.x + a.y * q) / den;
    r.y = (a.y - a.x * q) / den;
  } else {
    R q = b.x / b.y, den = b.y + b.x * q;
    r.x = (a.x * q + a.y) / den;
    r.y = (a.y * q - a.x) / den;
  }
  return r;
}
template <class T>
__device__ __forceinline__ T cfma(T a, T b, T c) {  // c + a*b
  T r;
  r.x = c.x + a.x * b.x - a.y * b.y;
  r.y = c.y + a.x * b.y + a.y * b.x;
  return r;
}
template <class T>
__device__ __forceinline__ T cfms(T a, T b, T c) {  // c - a*b
  T r;
  r.x = c.x - a.x * b.x + a.y * b.y;
  r.y = c.y - a.x * b.y - a.y * b.x;
  return r;
}
template <class T>
__device__ __forceinline__ typename CplxOf<T>::real cabsT(T a) {
  return hypot(a.x, a.y);
}

}  // namespace aifmm
