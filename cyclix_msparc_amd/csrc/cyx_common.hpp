// Shared device types and helpers for the Cyclix CheFSI / AAR library (gfx950).
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <cstdio>
#include <string>

namespace cyx {

// ---------------------------------------------------------------- scalars
template <class R>
struct alignas(2 * sizeof(R)) cplx {
  R x, y;
};

template <class T>
struct traits;
template <>
struct traits<double> {
  using real = double;
  static constexpr bool is_complex = false;
};
template <>
struct traits<float> {
  using real = float;
  static constexpr bool is_complex = false;
};
template <>
struct traits<cplx<double>> {
  using real = double;
  static constexpr bool is_complex = true;
};
template <>
struct traits<cplx<float>> {
  using real = float;
  static constexpr bool is_complex = true;
};

template <class T>
__host__ __device__ inline T zero() {
  return T{};
}
template <>
__host__ __device__ inline double zero<double>() {
  return 0.0;
}
template <>
__host__ __device__ inline float zero<float>() {
  return 0.0f;
}

// acc + a*x with a real
__device__ __forceinline__ double rfma(double a, double x, double acc) { return __builtin_fma(a, x, acc); }
__device__ __forceinline__ float rfma(float a, float x, float acc) { return __builtin_fmaf(a, x, acc); }
template <class R>
__device__ __forceinline__ cplx<R> rfma(R a, cplx<R> x, cplx<R> acc) {
  return cplx<R>{rfma(a, x.x, acc.x), rfma(a, x.y, acc.y)};
}

__device__ __forceinline__ double rmul(double a, double x) { return a * x; }
__device__ __forceinline__ float rmul(float a, float x) { return a * x; }
template <class R>
__device__ __forceinline__ cplx<R> rmul(R a, cplx<R> x) {
  return cplx<R>{a * x.x, a * x.y};
}

__device__ __forceinline__ double add(double a, double b) { return a + b; }
__device__ __forceinline__ float add(float a, float b) { return a + b; }
template <class R>
__device__ __forceinline__ cplx<R> add(cplx<R> a, cplx<R> b) {
  return cplx<R>{a.x + b.x, a.y + b.y};
}
__device__ __forceinline__ double sub(double a, double b) { return a - b; }
__device__ __forceinline__ float sub(float a, float b) { return a - b; }
template <class R>
__device__ __forceinline__ cplx<R> sub(cplx<R> a, cplx<R> b) {
  return cplx<R>{a.x - b.x, a.y - b.y};
}

// full product (phase * value); for real types phases are 1 and this is a*b
__device__ __forceinline__ double cmul(double a, double b) { return a * b; }
__device__ __forceinline__ float cmul(float a, float b) { return a * b; }
template <class R>
__device__ __forceinline__ cplx<R> cmul(cplx<R> a, cplx<R> b) {
  return cplx<R>{a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}
// acc + a*b (complex a, b)
__device__ __forceinline__ double cfma(double a, double b, double acc) { return __builtin_fma(a, b, acc); }
__device__ __forceinline__ float cfma(float a, float b, float acc) { return __builtin_fmaf(a, b, acc); }
template <class R>
__device__ __forceinline__ cplx<R> cfma(cplx<R> a, cplx<R> b, cplx<R> acc) {
  return cplx<R>{rfma(a.x, b.x, rfma(-a.y, b.y, acc.x)), rfma(a.x, b.y, rfma(a.y, b.x, acc.y))};
}
// acc + conj(a)*b
__device__ __forceinline__ double cjfma(double a, double b, double acc) { return __builtin_fma(a, b, acc); }
__device__ __forceinline__ float cjfma(float a, float b, float acc) { return __builtin_fmaf(a, b, acc); }
template <class R>
__device__ __forceinline__ cplx<R> cjfma(cplx<R> a, cplx<R> b, cplx<R> acc) {
  return cplx<R>{rfma(a.x, b.x, rfma(a.y, b.y, acc.x)), rfma(a.x, b.y, rfma(-a.y, b.x, acc.y))};
}

__device__ __forceinline__ double conjg(double a) { return a; }
__device__ __forceinline__ float conjg(float a) { return a; }
template <class R>
__device__ __forceinline__ cplx<R> conjg(cplx<R> a) {
  return cplx<R>{a.x, -a.y};
}
__device__ __forceinline__ double absq(double a) { return a * a; }
__device__ __forceinline__ float absq(float a) { return a * a; }
template <class R>
__device__ __forceinline__ R absq(cplx<R> a) {
  return a.x * a.x + a.y * a.y;
}

// FP64 accumulator type of T
template <class T>
struct acc_of {
  using type = double;
};
template <class R>
struct acc_of<cplx<R>> {
  using type = cplx<double>;
};

__device__ __forceinline__ double widen(double a) { return a; }
__device__ __forceinline__ double widen(float a) { return (double)a; }
__device__ __forceinline__ cplx<double> widen(cplx<double> a) { return a; }
__device__ __forceinline__ cplx<double> widen(cplx<float> a) { return cplx<double>{(double)a.x, (double)a.y}; }

template <class T>
__device__ __forceinline__ T narrow(typename acc_of<T>::type a);
template <>
__device__ __forceinline__ double narrow<double>(double a) {
  return a;
}
template <>
__device__ __forceinline__ float narrow<float>(double a) {
  return (float)a;
}
template <>
__device__ __forceinline__ cplx<double> narrow<cplx<double>>(cplx<double> a) {
  return a;
}
template <>
__device__ __forceinline__ cplx<float> narrow<cplx<float>>(cplx<double> a) {
  return cplx<float>{(float)a.x, (float)a.y};
}

// host-side phase e^{i phi} in type T (real types: only phi == 0 is valid)
template <class T>
inline T phase_of(double c, double s);
template <>
inline double phase_of<double>(double c, double) {
  return c;
}
template <>
inline float phase_of<float>(double c, double) {
  return (float)c;
}
template <>
inline cplx<double> phase_of<cplx<double>>(double c, double s) {
  return cplx<double>{c, s};
}
template <>
inline cplx<float> phase_of<cplx<float>>(double c, double s) {
  return cplx<float>{(float)c, (float)s};
}

// ---------------------------------------------------------------- errors
void set_error(const char* fmt, ...);

#define CYX_HIP(call)                                                                  \
  do {                                                                                 \
    hipError_t e_ = (call);                                                            \
    if (e_ != hipSuccess) {                                                            \
      ::cyx::set_error("%s:%d %s: %s", __FILE__, __LINE__, #call, hipGetErrorString(e_)); \
      return CYX_ERR_HIP;                                                              \
    }                                                                                  \
  } while (0)

#define CYX_CHECK(cond, ...)      \
  do {                            \
    if (!(cond)) {                \
      ::cyx::set_error(__VA_ARGS__); \
      return CYX_ERR_ARG;         \
    }                             \
  } while (0)

inline int ceil_div(long a, long b) { return (int)((a + b - 1) / b); }

}  // namespace cyx
