// Common definitions for the HPS MI355X (gfx950) kernels.
//
// Scalar types: double, float, and interleaved complex (re, im) of each.
// All matrices are row-major with an explicit leading dimension (row stride)
// and a batch stride; a batch stride of 0 broadcasts one matrix to the batch.
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>
#include <cstddef>

namespace hps {

struct c64 {
  float re, im;
};
struct c128 {
  double re, im;
};

template <typename T> struct real_of { using type = T; };
template <> struct real_of<c64> { using type = float; };
template <> struct real_of<c128> { using type = double; };

template <typename T> struct is_cplx { static constexpr bool value = false; };
template <> struct is_cplx<c64> { static constexpr bool value = true; };
template <> struct is_cplx<c128> { static constexpr bool value = true; };

// ---- scalar helpers (device + host) ---------------------------------------
__host__ __device__ inline double s_re(double x) { return x; }
__host__ __device__ inline float s_re(float x) { return x; }
__host__ __device__ inline double s_re(c128 x) { return x.re; }
__host__ __device__ inline float s_re(c64 x) { return x.re; }
__host__ __device__ inline double s_im(double) { return 0.0; }
__host__ __device__ inline float s_im(float) { return 0.0f; }
__host__ __device__ inline double s_im(c128 x) { return x.im; }
__host__ __device__ inline float s_im(c64 x) { return x.im; }

template <typename T> __host__ __device__ inline T s_zero() { return T(0); }
template <> __host__ __device__ inline c64 s_zero<c64>() { return c64{0.f, 0.f}; }
template <> __host__ __device__ inline c128 s_zero<c128>() { return c128{0.0, 0.0}; }

template <typename T> __host__ __device__ inline T s_one() { return T(1); }
template <> __host__ __device__ inline c64 s_one<c64>() { return c64{1.f, 0.f}; }
template <> __host__ __device__ inline c128 s_one<c128>() { return c128{1.0, 0.0}; }

__host__ __device__ inline double s_add(double a, double b) { return a + b; }
__host__ __device__ inline float s_add(float a, float b) { return a + b; }
__host__ __device__ inline c128 s_add(c128 a, c128 b) { return c128{a.re + b.re, a.im + b.im}; }
__host__ __device__ inline c64 s_add(c64 a, c64 b) { return c64{a.re + b.re, a.im + b.im}; }

__host__ __device__ inline double s_sub(double a, double b) { return a - b; }
__host__ __device__ inline float s_sub(float a, float b) { return a - b; }
__host__ __device__ inline c128 s_sub(c128 a, c128 b) { return c128{a.re - b.re, a.im - b.im}; }
__host__ __device__ inline c64 s_sub(c64 a, c64 b) { return c64{a.re - b.re, a.im - b.im}; }

__host__ __device__ inline double s_mul(double a, double b) { return a * b; }
__host__ __device__ inline float s_mul(float a, float b) { return a * b; }
__host__ __device__ inline c128 s_mul(c128 a, c128 b) {
  return c128{a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}
__host__ __device__ inline c64 s_mul(c64 a, c64 b) {
  return c64{a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}

// a - b*c
template <typename T> __host__ __device__ inline T s_fnms(T a, T b, T c) { return s_sub(a, s_mul(b, c)); }
// a*b + c
template <typename T> __host__ __device__ inline T s_fma(T a, T b, T c) { return s_add(s_mul(a, b), c); }
// complex forms as four fused multiply-adds each (the generic form costs a
// separate multiply and add per component)
template <typename C, typename R>
__host__ __device__ inline C c_fnms(C a, C b, C c) {
  const R re = fma(-b.re, c.re, fma(b.im, c.im, a.re));
  const R im = fma(-b.re, c.im, fma(-b.im, c.re, a.im));
  return C{re, im};
}
template <typename C, typename R>
__host__ __device__ inline C c_fma(C a, C b, C c) {
  const R re = fma(a.re, b.re, fma(-a.im, b.im, c.re));
  const R im = fma(a.re, b.im, fma(a.im, b.re, c.im));
  return C{re, im};
}
template <> __host__ __device__ inline c128 s_fnms<c128>(c128 a, c128 b, c128 c) { return c_fnms<c128, double>(a, b, c); }
template <> __host__ __device__ inline c64 s_fnms<c64>(c64 a, c64 b, c64 c) { return c_fnms<c64, float>(a, b, c); }
template <> __host__ __device__ inline c128 s_fma<c128>(c128 a, c128 b, c128 c) { return c_fma<c128, double>(a, b, c); }
template <> __host__ __device__ inline c64 s_fma<c64>(c64 a, c64 b, c64 c) { return c_fma<c64, float>(a, b, c); }

__host__ __device__ inline double s_scale(double a, double s) { return a * s; }
__host__ __device__ inline float s_scale(float a, double s) { return a * (float)s; }
__host__ __device__ inline c128 s_scale(c128 a, double s) { return c128{a.re * s, a.im * s}; }
__host__ __device__ inline c64 s_scale(c64 a, double s) { return c64{a.re * (float)s, a.im * (float)s}; }

__host__ __device__ inline double s_neg(double a) { return -a; }
__host__ __device__ inline float s_neg(float a) { return -a; }
__host__ __device__ inline c128 s_neg(c128 a) { return c128{-a.re, -a.im}; }
__host__ __device__ inline c64 s_neg(c64 a) { return c64{-a.re, -a.im}; }

// 1 / a
__host__ __device__ inline double s_inv(double a) { return 1.0 / a; }
__host__ __device__ inline float s_inv(float a) { return 1.0f / a; }
__host__ __device__ inline c128 s_inv(c128 a) {
  // Smith's algorithm for a robust complex reciprocal.
  if (fabs(a.re) >= fabs(a.im)) {
    double r = a.im / a.re, d = a.re + a.im * r;
    return c128{1.0 / d, -r / d};
  } else {
    double r = a.re / a.im, d = a.re * r + a.im;
    return c128{r / d, -1.0 / d};
  }
}
__host__ __device__ inline c64 s_inv(c64 a) {
  if (fabsf(a.re) >= fabsf(a.im)) {
    float r = a.im / a.re, d = a.re + a.im * r;
    return c64{1.0f / d, -r / d};
  } else {
    float r = a.re / a.im, d = a.re * r + a.im;
    return c64{r / d, -1.0f / d};
  }
}

// pivot magnitude |re| + |im| (the BLAS i*amax convention)
__host__ __device__ inline double s_abs1(double a) { return fabs(a); }
__host__ __device__ inline float s_abs1(float a) { return fabsf(a); }
__host__ __device__ inline double s_abs1(c128 a) { return fabs(a.re) + fabs(a.im); }
__host__ __device__ inline float s_abs1(c64 a) { return fabsf(a.re) + fabsf(a.im); }

__host__ __device__ inline bool s_is_zero(double a) { return a == 0.0; }
__host__ __device__ inline bool s_is_zero(float a) { return a == 0.0f; }
__host__ __device__ inline bool s_is_zero(c128 a) { return a.re == 0.0 && a.im == 0.0; }
__host__ __device__ inline bool s_is_zero(c64 a) { return a.re == 0.0f && a.im == 0.0f; }

// squared magnitude, accumulated in double
__host__ __device__ inline double s_norm2(double a) { return a * a; }
__host__ __device__ inline double s_norm2(float a) { return (double)a * a; }
__host__ __device__ inline double s_norm2(c128 a) { return a.re * a.re + a.im * a.im; }
__host__ __device__ inline double s_norm2(c64 a) { return (double)a.re * a.re + (double)a.im * a.im; }

// conversion between the four scalar types (complex -> real keeps the real part)
template <typename To> struct cvt_to;
template <> struct cvt_to<double> {
  template <typename F> __host__ __device__ static double from(F x) { return (double)s_re(x); }
};
template <> struct cvt_to<float> {
  template <typename F> __host__ __device__ static float from(F x) { return (float)s_re(x); }
};
template <> struct cvt_to<c128> {
  template <typename F> __host__ __device__ static c128 from(F x) { return c128{(double)s_re(x), (double)s_im(x)}; }
};
template <> struct cvt_to<c64> {
  template <typename F> __host__ __device__ static c64 from(F x) { return c64{(float)s_re(x), (float)s_im(x)}; }
};
template <typename To, typename F> __host__ __device__ inline To s_cvt(F x) { return cvt_to<To>::from(x); }

// Precision codes shared with the C ABI (include/hps_mi355x.h).
enum Dtype : int { HPS_F64 = 0, HPS_F32 = 1, HPS_C128 = 2, HPS_C64 = 3 };

inline size_t dtype_size(int dt) {
  switch (dt) {
    case HPS_F64: return 8;
    case HPS_F32: return 4;
    case HPS_C128: return 16;
    case HPS_C64: return 8;
  }
  return 0;
}

template <typename T> inline const char* tname();
template <> inline const char* tname<double>() { return "f64"; }
template <> inline const char* tname<float>() { return "f32"; }
template <> inline const char* tname<c128>() { return "c128"; }
template <> inline const char* tname<c64>() { return "c64"; }
template <typename T> constexpr double flop_mult() { return is_cplx<T>::value ? 4.0 : 1.0; }

inline unsigned ceil_div(long a, long b) { return (unsigned)((a + b - 1) / b); }

}  // namespace hps
