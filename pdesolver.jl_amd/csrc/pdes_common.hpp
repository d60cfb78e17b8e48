// pdes_common.hpp -- shared host/device helpers of the gfx950 library.
#pragma once
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <string>
#include <type_traits>
#include <utility>

#include "../../include/pdes.h"

namespace pdes {

void set_error(const char* fmt, ...);

#define PDES_HIP(call)                                                          \
  do {                                                                          \
    hipError_t _e = (call);                                                     \
    if (_e != hipSuccess) {                                                     \
      ::pdes::set_error("%s:%d %s: %s", __FILE__, __LINE__, #call,              \
                        hipGetErrorString(_e));                                 \
      return PDES_ERR_HIP;                                                      \
    }                                                                           \
  } while (0)

#define PDES_CHECK(cond, code, ...)                                             \
  do {                                                                          \
    if (!(cond)) {                                                              \
      ::pdes::set_error(__VA_ARGS__);                                           \
      return (code);                                                            \
    }                                                                           \
  } while (0)

// ---------------------------------------------------------------------------
// Forward-mode dual numbers.  With h -> 0 the complex step of the reference
// (q + i h v, src/interface2.jl:454-498) propagates exactly these first-order
// derivatives; every branch of the flux functions tests the real part, as
// the reference's complex-safe helpers do (src/Utils/complexify.jl).
// ---------------------------------------------------------------------------
template <int N>
struct dual {
  double v;
  double d[N];
  __host__ __device__ dual() {}
  __host__ __device__ dual(double x) : v(x) {
#pragma unroll
    for (int i = 0; i < N; ++i) d[i] = 0.0;
  }
};

template <int N> __device__ __forceinline__ double re(const dual<N>& a) { return a.v; }
__device__ __forceinline__ double re(double a) { return a; }

// The batch that workgroup g of nb works on, so that each XCD walks a
// contiguous eighth of the batches: the hardware deals workgroups to the 8
// XCDs round-robin (workgroup g runs on XCD g % 8), and XCD x takes batches
// [x q + min(x, r), ...) of nb = 8 q + r.  Neighbouring batches, whose data
// overlap (an element's face neighbours), are then read through the same L2.
// A bijection on [0, nb).
__device__ __forceinline__ int64_t xcd_contiguous_batch(int64_t g, int64_t nb) {
  const int64_t q = nb / 8, r = nb % 8, x = g % 8;
  return x * q + (x < r ? x : r) + g / 8;
}

// 1/x from the hardware reciprocal estimate (v_rcp_f64) refined by two
// Newton-Raphson steps r <- r + r (1 - x r), each of which squares the
// relative error.  Within about one ulp, not correctly rounded; meant for the
// well-scaled positive sums of the flux functions, far from the exponent
// limits where an IEEE division needs its scaling steps.
__device__ __forceinline__ double drcp(double x) {
  double r = __builtin_amdgcn_rcp(x);
  double e = fma(-x, r, 1.0);
  r = fma(r, e, r);
  e = fma(-x, r, 1.0);
  return fma(r, e, r);
}
// sqrt(x) for positive normal x: the hardware reciprocal square root estimate
// (v_rsq_f64) refined by the coupled Newton iteration for g ~ sqrt(x) and
// h ~ 1/(2 sqrt(x)) (Goldschmidt's square root iteration, as in P. Markstein,
// "IA-64 and Elementary Functions", 2000), ending with two corrections
// g += (x - g^2) h.  Within about one ulp.  Zero, subnormal, infinite and
// negative arguments are not handled: the flux functions take square roots of
// positive physical quantities only, and non-physical states are flagged by
// the positivity checks.
__device__ __forceinline__ double dsqrt(double x) {
  const double y = __builtin_amdgcn_rsq(x);
  double g = x * y, h = 0.5 * y;
  const double r = fma(-g, h, 0.5);
  g = fma(g, r, g);
  h = fma(h, r, h);
  double d = fma(-g, g, x);
  g = fma(d, h, g);
  d = fma(-g, g, x);
  return fma(d, h, g);
}

#define DUAL_BINOP(OP, VEXPR, DEXPR)                                              \
  template <int N> __device__ __forceinline__ dual<N> operator OP(const dual<N>& a, const dual<N>& b) { \
    dual<N> r; r.v = VEXPR;                                                       \
    _Pragma("unroll") for (int i = 0; i < N; ++i) r.d[i] = DEXPR;                 \
    return r; }
DUAL_BINOP(+, a.v + b.v, a.d[i] + b.d[i])
DUAL_BINOP(-, a.v - b.v, a.d[i] - b.d[i])
DUAL_BINOP(*, a.v * b.v, a.d[i] * b.v + a.v * b.d[i])
#undef DUAL_BINOP

template <int N> __device__ __forceinline__ dual<N> operator/(const dual<N>& a, const dual<N>& b) {
  dual<N> r;
  double inv = 1.0 / b.v;
  r.v = a.v * inv;
#pragma unroll
  for (int i = 0; i < N; ++i) r.d[i] = (a.d[i] - r.v * b.d[i]) * inv;
  return r;
}
template <int N> __device__ __forceinline__ dual<N> operator-(const dual<N>& a) {
  dual<N> r; r.v = -a.v;
#pragma unroll
  for (int i = 0; i < N; ++i) r.d[i] = -a.d[i];
  return r;
}
// mixed dual/double
template <int N> __device__ __forceinline__ dual<N> operator+(const dual<N>& a, double b) { dual<N> r = a; r.v += b; return r; }
template <int N> __device__ __forceinline__ dual<N> operator+(double b, const dual<N>& a) { dual<N> r = a; r.v += b; return r; }
template <int N> __device__ __forceinline__ dual<N> operator-(const dual<N>& a, double b) { dual<N> r = a; r.v -= b; return r; }
template <int N> __device__ __forceinline__ dual<N> operator-(double b, const dual<N>& a) {
  dual<N> r; r.v = b - a.v;
#pragma unroll
  for (int i = 0; i < N; ++i) r.d[i] = -a.d[i];
  return r;
}
template <int N> __device__ __forceinline__ dual<N> operator*(const dual<N>& a, double b) {
  dual<N> r; r.v = a.v * b;
#pragma unroll
  for (int i = 0; i < N; ++i) r.d[i] = a.d[i] * b;
  return r;
}
template <int N> __device__ __forceinline__ dual<N> operator*(double b, const dual<N>& a) { return a * b; }
template <int N> __device__ __forceinline__ dual<N> operator/(const dual<N>& a, double b) {
  dual<N> r; double inv = 1.0 / b; r.v = a.v / b;
#pragma unroll
  for (int i = 0; i < N; ++i) r.d[i] = a.d[i] * inv;
  return r;
}
template <int N> __device__ __forceinline__ dual<N> operator/(double a, const dual<N>& b) {
  dual<N> r; double inv = 1.0 / b.v; r.v = a * inv;
#pragma unroll
  for (int i = 0; i < N; ++i) r.d[i] = -r.v * b.d[i] * inv;
  return r;
}
template <int N> __device__ __forceinline__ dual<N>& operator+=(dual<N>& a, const dual<N>& b) { a = a + b; return a; }
template <int N> __device__ __forceinline__ dual<N>& operator-=(dual<N>& a, const dual<N>& b) { a = a - b; return a; }

template <int N> __device__ __forceinline__ dual<N> dsqrt(const dual<N>& a) {
  dual<N> r; r.v = dsqrt(a.v); double f = 0.5 * drcp(r.v);
#pragma unroll
  for (int i = 0; i < N; ++i) r.d[i] = a.d[i] * f;
  return r;
}
template <int N> __device__ __forceinline__ dual<N> dlog(const dual<N>& a) {
  dual<N> r; r.v = log(a.v); double f = 1.0 / a.v;
#pragma unroll
  for (int i = 0; i < N; ++i) r.d[i] = a.d[i] * f;
  return r;
}
template <int N> __device__ __forceinline__ dual<N> dpow(const dual<N>& a, double g) {
  dual<N> r; r.v = pow(a.v, g); double f = g * r.v / a.v;
#pragma unroll
  for (int i = 0; i < N; ++i) r.d[i] = a.d[i] * f;
  return r;
}
template <int N> __device__ __forceinline__ dual<N> drcp(const dual<N>& a) {
  dual<N> r;
  r.v = drcp(a.v);
  const double f = -r.v * r.v;
#pragma unroll
  for (int i = 0; i < N; ++i) r.d[i] = a.d[i] * f;
  return r;
}
__device__ __forceinline__ double dlog(double a) { return log(a); }
__device__ __forceinline__ double dpow(double a, double g) { return pow(a, g); }
__device__ __forceinline__ double dexp(double a) { return exp(a); }
template <int N> __device__ __forceinline__ dual<N> dexp(const dual<N>& a) {
  dual<N> r;
  r.v = exp(a.v);
#pragma unroll
  for (int k = 0; k < N; ++k) r.d[k] = r.v * a.d[k];
  return r;
}

// absvalue (src/Utils/complexify.jl:25-45): sign flip by the real part
__device__ __forceinline__ double absvalue(double x) { return fabs(x); }
template <int N> __device__ __forceinline__ dual<N> absvalue(const dual<N>& x) { return x.v < 0 ? -x : x; }
// Julia max(x, y) = isless(y, x) ? x : y, ordered by the real part
template <class T> __device__ __forceinline__ T jmax(const T& x, const T& y) { return re(y) < re(x) ? x : y; }
// absvalue3 (src/Utils/complexify.jl:161-172)
template <class T> __device__ __forceinline__ T absvalue3(const T& val) {
  const double delta = 1e-7;
  T val1 = absvalue(val);
  if (re(val1) > delta) return val1;
  return ((val * val) / delta + delta) / 2.0;
}

inline int64_t ceil_div(int64_t a, int64_t b) { return (a + b - 1) / b; }

}  // namespace pdes
