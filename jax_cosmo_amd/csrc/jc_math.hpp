// Scalar math and forward-mode dual numbers for the Limber C_ell kernels.
//
// Every physics routine of the library is written once, as a template on its
// scalar type S:
//   float / double          -- plain values (the forward C_ell path),
//   Dual<float|double, N>   -- a value plus N directional derivatives.
// The derivative kernels (JVP and the adjoint/VJP) instantiate the very same
// code with Dual types, so the derivative is exactly that of the discrete
// algorithm the forward path runs (frozen interpolation indices, zero
// derivative through an active clip), which is what reverse-mode AD of the
// reference computes.
//
// float uses the gfx950 hardware transcendental instructions (v_exp_f32,
// v_log_f32, v_rcp_f32, v_sqrt_f32: ~1 ulp); double uses the precise device
// libm.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define JC_HD __host__ __device__ __forceinline__
#define JC_D __device__ __forceinline__

namespace jm {

// ------------------------------------------------------------ float
JC_D float exp(float x) { return __builtin_amdgcn_exp2f(x * 1.4426950408889634f); }
JC_D float log(float x) { return __builtin_amdgcn_logf(x) * 0.6931471805599453f; }
JC_D float exp2(float x) { return __builtin_amdgcn_exp2f(x); }
JC_D float log2(float x) { return __builtin_amdgcn_logf(x); }
JC_D float rcp(float x) { return __builtin_amdgcn_rcpf(x); }
JC_D float sqrt(float x) { return __builtin_amdgcn_sqrtf(x); }
JC_D float div(float a, float b) { return a * __builtin_amdgcn_rcpf(b); }
// sin / cos through the hardware v_sin_f32 / v_cos_f32, whose input is in
// revolutions: x / (2 pi) is formed with an error-compensated product and its
// integer part removed, so the reduction stays accurate to ~1e-7 revolutions
// for |x| up to ~1e6.
constexpr float INV2PI_HI = 0.15915494f;
constexpr float INV2PI_LO = (float)(0.15915494309189535 - (double)0.15915494f);
JC_D float revolutions(float x) {
  const float r = x * INV2PI_HI;
  const float e = __builtin_fmaf(x, INV2PI_HI, -r) + x * INV2PI_LO;
  return (r - __builtin_rintf(r)) + e;
}
JC_D float sin(float x) { return __builtin_amdgcn_sinf(revolutions(x)); }
JC_D float cos(float x) { return __builtin_amdgcn_cosf(revolutions(x)); }
JC_D float fabs(float x) { return __builtin_fabsf(x); }
JC_D float pow(float x, float y) { return jm::exp(y * jm::log(x)); }

// ------------------------------------------------------------ double
JC_D double exp(double x) { return ::exp(x); }
JC_D double log(double x) { return ::log(x); }
JC_D double exp2(double x) { return ::exp2(x); }
JC_D double log2(double x) { return ::log2(x); }
JC_D double rcp(double x) { return 1.0 / x; }
JC_D double sqrt(double x) { return ::sqrt(x); }
JC_D double div(double a, double b) { return a / b; }
JC_D double sin(double x) { return ::sin(x); }
JC_D double cos(double x) { return ::cos(x); }
JC_D double fabs(double x) { return __builtin_fabs(x); }
JC_D double pow(double x, double y) { return ::pow(x, y); }

// a / b: hardware reciprocal for float (about 1 ulp), exact division for double
JC_D float fdiv(float a, float b) { return a * __builtin_amdgcn_rcpf(b); }
JC_D double fdiv(double a, double b) { return a / b; }

// e^x to float accuracy (about 2e-7 relative, |x| < 80) for values that are
// stored as floats: t = x log2(e) in FP64 is split into n + f with |f| <= 1/2
// and 2^f comes from the hardware v_exp_f32.  A float argument uses exp.
JC_D double expf_acc(double x) {
  const double t = x * 1.4426950408889634;
  const double n = __builtin_rint(t);
  return __builtin_ldexp((double)__builtin_amdgcn_exp2f((float)(t - n)), (int)n);
}
JC_D float expf_acc(float x) { return exp(x); }

// ln x (x > 0, normal) to float accuracy (about 1e-7 absolute) for values
// that are stored as floats or only feed float-accurate exponentials:
// x = m 2^e with m in [1/sqrt 2, sqrt 2) is split exactly in FP64, and log2 m
// comes from the hardware v_log_f32 of the float-rounded m.  A float
// argument uses log.
JC_D double logf_acc(double x) {
  int e = 0;
  double m = __builtin_frexp(x, &e);  // m in [1/2, 1)
  if (m < 0.70710678118654752) {
    m *= 2.0;
    e -= 1;
  }
  return ((double)e + (double)__builtin_amdgcn_logf((float)m)) * 0.69314718055994531;
}
JC_D float logf_acc(float x) { return log(x); }

}  // namespace jm

// ------------------------------------------------------------ dual numbers
template <typename T, int N>
struct Dual {
  T v;
  T d[N];
};

template <typename S>
struct Base { using type = S; static constexpr int n = 0; };
template <typename T, int N>
struct Base<Dual<T, N>> { using type = T; static constexpr int n = N; };

// non-deduced context, so Dual<float,N> op double-literal converts the literal
template <typename T>
struct Id { using type = T; };

// value access for both plain and dual scalars
JC_HD float val(float x) { return x; }
JC_HD double val(double x) { return x; }
template <typename T, int N>
JC_HD T val(const Dual<T, N>& x) { return x.v; }

// construction from a plain constant
template <typename S>
struct Make {
  JC_HD static S c(double x) { return (S)x; }
};
template <typename T, int N>
struct Make<Dual<T, N>> {
  JC_HD static Dual<T, N> c(double x) {
    Dual<T, N> r;
    r.v = (T)x;
#pragma unroll
    for (int i = 0; i < N; ++i) r.d[i] = (T)0;
    return r;
  }
};

// precision conversion (double <-> float), plain and dual
template <typename To, typename From>
struct Conv {
  JC_HD static To c(const From& x) { return (To)x; }
};
template <typename T1, typename T2, int N>
struct Conv<Dual<T1, N>, Dual<T2, N>> {
  JC_HD static Dual<T1, N> c(const Dual<T2, N>& x) {
    Dual<T1, N> r;
    r.v = (T1)x.v;
#pragma unroll
    for (int i = 0; i < N; ++i) r.d[i] = (T1)x.d[i];
    return r;
  }
};
template <typename To, typename From>
JC_HD To conv(const From& x) { return Conv<To, From>::c(x); }

#define DUAL_T template <typename T, int N>
#define DL Dual<T, N>

DUAL_T JC_D DL operator+(const DL& a, const DL& b) {
  DL r; r.v = a.v + b.v;
#pragma unroll
  for (int i = 0; i < N; ++i) r.d[i] = a.d[i] + b.d[i];
  return r;
}
DUAL_T JC_D DL operator-(const DL& a, const DL& b) {
  DL r; r.v = a.v - b.v;
#pragma unroll
  for (int i = 0; i < N; ++i) r.d[i] = a.d[i] - b.d[i];
  return r;
}
DUAL_T JC_D DL operator-(const DL& a) {
  DL r; r.v = -a.v;
#pragma unroll
  for (int i = 0; i < N; ++i) r.d[i] = -a.d[i];
  return r;
}
DUAL_T JC_D DL operator*(const DL& a, const DL& b) {
  DL r; r.v = a.v * b.v;
#pragma unroll
  for (int i = 0; i < N; ++i) r.d[i] = a.v * b.d[i] + a.d[i] * b.v;
  return r;
}
DUAL_T JC_D DL operator/(const DL& a, const DL& b) {
  T ib = jm::rcp(b.v);
  DL r; r.v = a.v * ib;
#pragma unroll
  for (int i = 0; i < N; ++i) r.d[i] = (a.d[i] - r.v * b.d[i]) * ib;
  return r;
}
// mixed with plain constants (T or convertible)
DUAL_T JC_D DL operator+(const DL& a, typename Id<T>::type b) { DL r = a; r.v += b; return r; }
DUAL_T JC_D DL operator+(typename Id<T>::type b, const DL& a) { DL r = a; r.v += b; return r; }
DUAL_T JC_D DL operator-(const DL& a, typename Id<T>::type b) { DL r = a; r.v -= b; return r; }
DUAL_T JC_D DL operator-(typename Id<T>::type b, const DL& a) {
  DL r; r.v = b - a.v;
#pragma unroll
  for (int i = 0; i < N; ++i) r.d[i] = -a.d[i];
  return r;
}
DUAL_T JC_D DL operator*(const DL& a, typename Id<T>::type b) {
  DL r; r.v = a.v * b;
#pragma unroll
  for (int i = 0; i < N; ++i) r.d[i] = a.d[i] * b;
  return r;
}
DUAL_T JC_D DL operator*(typename Id<T>::type b, const DL& a) { return a * b; }
DUAL_T JC_D DL operator/(const DL& a, typename Id<T>::type b) {
  T ib = jm::rcp(b);
  return a * ib;
}
DUAL_T JC_D DL operator/(typename Id<T>::type a, const DL& b) {
  T ib = jm::rcp(b.v);
  DL r; r.v = a * ib;
  T s = -r.v * ib;
#pragma unroll
  for (int i = 0; i < N; ++i) r.d[i] = s * b.d[i];
  return r;
}
DUAL_T JC_D DL& operator+=(DL& a, const DL& b) { a = a + b; return a; }
DUAL_T JC_D DL& operator-=(DL& a, const DL& b) { a = a - b; return a; }
DUAL_T JC_D DL& operator*=(DL& a, const DL& b) { a = a * b; return a; }

namespace jm {
DUAL_T JC_D DL exp(const DL& a) {
  DL r; r.v = jm::exp(a.v);
#pragma unroll
  for (int i = 0; i < N; ++i) r.d[i] = r.v * a.d[i];
  return r;
}
DUAL_T JC_D DL exp2(const DL& a) {
  DL r; r.v = jm::exp2(a.v);
  const T s = r.v * T(0.6931471805599453);
#pragma unroll
  for (int i = 0; i < N; ++i) r.d[i] = s * a.d[i];
  return r;
}
DUAL_T JC_D DL log2(const DL& a) {
  DL r; r.v = jm::log2(a.v);
  const T ia = jm::rcp(a.v) * T(1.4426950408889634);
#pragma unroll
  for (int i = 0; i < N; ++i) r.d[i] = a.d[i] * ia;
  return r;
}
DUAL_T JC_D DL expf_acc(const DL& a) {
  DL r; r.v = jm::expf_acc(a.v);
#pragma unroll
  for (int i = 0; i < N; ++i) r.d[i] = r.v * a.d[i];
  return r;
}
DUAL_T JC_D DL logf_acc(const DL& a) {
  DL r; r.v = jm::logf_acc(a.v);
  T ia = jm::rcp(a.v);
#pragma unroll
  for (int i = 0; i < N; ++i) r.d[i] = a.d[i] * ia;
  return r;
}
DUAL_T JC_D DL log(const DL& a) {
  DL r; r.v = jm::log(a.v);
  T ia = jm::rcp(a.v);
#pragma unroll
  for (int i = 0; i < N; ++i) r.d[i] = a.d[i] * ia;
  return r;
}
DUAL_T JC_D DL sqrt(const DL& a) {
  DL r; r.v = jm::sqrt(a.v);
  T s = (T)0.5 * jm::rcp(r.v);
#pragma unroll
  for (int i = 0; i < N; ++i) r.d[i] = a.d[i] * s;
  return r;
}
DUAL_T JC_D DL rcp(const DL& a) { return (T)1 / a; }
DUAL_T JC_D DL div(const DL& a, const DL& b) { return a / b; }
DUAL_T JC_D DL fdiv(const DL& a, const DL& b) { return a / b; }
DUAL_T JC_D DL fdiv(typename Id<T>::type a, const DL& b) { return a / b; }
DUAL_T JC_D DL fdiv(const DL& a, typename Id<T>::type b) { return a * jm::rcp(b); }
DUAL_T JC_D DL sin(const DL& a) {
  DL r; r.v = jm::sin(a.v);
  T c = jm::cos(a.v);
#pragma unroll
  for (int i = 0; i < N; ++i) r.d[i] = a.d[i] * c;
  return r;
}
DUAL_T JC_D DL cos(const DL& a) {
  DL r; r.v = jm::cos(a.v);
  T s = -jm::sin(a.v);
#pragma unroll
  for (int i = 0; i < N; ++i) r.d[i] = a.d[i] * s;
  return r;
}
DUAL_T JC_D DL fabs(const DL& a) {
  // d|x| = sign(x) dx; JAX uses sign(0) = 0 at the kink
  T s = a.v > (T)0 ? (T)1 : (a.v < (T)0 ? (T)-1 : (T)0);
  DL r; r.v = jm::fabs(a.v);
#pragma unroll
  for (int i = 0; i < N; ++i) r.d[i] = a.d[i] * s;
  return r;
}
// x^c for a constant exponent
DUAL_T JC_D DL pow(const DL& a, typename Id<T>::type c) {
  DL r; r.v = jm::pow(a.v, c);
  T s = c * r.v * jm::rcp(a.v);
#pragma unroll
  for (int i = 0; i < N; ++i) r.d[i] = a.d[i] * s;
  return r;
}
// x^y for a dual exponent: exp(y ln x)
DUAL_T JC_D DL pow(const DL& a, const DL& y) { return jm::exp(y * jm::log(a)); }
}  // namespace jm

// power with a plain-constant exponent for plain scalars too
namespace jm {
JC_D float powc(float x, float c) { return jm::pow(x, c); }
JC_D double powc(double x, double c) { return jm::pow(x, c); }
DUAL_T JC_D DL powc(const DL& x, typename Id<T>::type c) { return jm::pow(x, c); }
}  // namespace jm

// maximum(x, c) / minimum(x, c) with a constant bound, JAX semantics: the
// derivative passes where x is strictly on the kept side and is halved at a
// tie (jax.numpy.maximum splits the cotangent evenly).
JC_D float maxc(float x, float c) { return x > c ? x : c; }
JC_D double maxc(double x, double c) { return x > c ? x : c; }
JC_D float minc(float x, float c) { return x < c ? x : c; }
JC_D double minc(double x, double c) { return x < c ? x : c; }
DUAL_T JC_D DL maxc(const DL& x, typename Id<T>::type c) {
  DL r;
  T s = x.v > c ? (T)1 : (x.v == c ? (T)0.5 : (T)0);
  r.v = x.v > c ? x.v : c;
#pragma unroll
  for (int i = 0; i < N; ++i) r.d[i] = x.d[i] * s;
  return r;
}
DUAL_T JC_D DL minc(const DL& x, typename Id<T>::type c) {
  DL r;
  T s = x.v < c ? (T)1 : (x.v == c ? (T)0.5 : (T)0);
  r.v = x.v < c ? x.v : c;
#pragma unroll
  for (int i = 0; i < N; ++i) r.d[i] = x.d[i] * s;
  return r;
}
// maximum of two variables (chi' - chi clip uses a constant 0 instead)
#undef DUAL_T
#undef DL

// ------------------------------------------------------------ scalar traits
template <int ND>
struct Types {
  using D = Dual<double, ND>;
  using F = Dual<float, ND>;
};
template <>
struct Types<0> {
  using D = double;
  using F = float;
};

// number of floats a scalar of type F occupies in memory
template <typename S>
struct Width { static constexpr int value = 1; };
template <typename T, int N>
struct Width<Dual<T, N>> { static constexpr int value = N + 1; };

// memory load/store of a (possibly dual) float scalar at p[0..W)
// (the strided forms place component i at p[i * stride])
template <typename S>
struct Mem {
  JC_D static S ld(const float* p) { return p[0]; }
  JC_D static void st(float* p, const S& x) { p[0] = x; }
  JC_D static S ld(const float* p, int) { return p[0]; }
  JC_D static void st(float* p, const S& x, int) { p[0] = x; }
};
template <int N>
struct Mem<Dual<float, N>> {
  JC_D static Dual<float, N> ld(const float* p) {
    Dual<float, N> r;
    r.v = p[0];
#pragma unroll
    for (int i = 0; i < N; ++i) r.d[i] = p[1 + i];
    return r;
  }
  JC_D static void st(float* p, const Dual<float, N>& x) {
    p[0] = x.v;
#pragma unroll
    for (int i = 0; i < N; ++i) p[1 + i] = x.d[i];
  }
  JC_D static Dual<float, N> ld(const float* p, int stride) {
    Dual<float, N> r;
    r.v = p[0];
#pragma unroll
    for (int i = 0; i < N; ++i) r.d[i] = p[(1 + i) * stride];
    return r;
  }
  JC_D static void st(float* p, const Dual<float, N>& x, int stride) {
    p[0] = x.v;
#pragma unroll
    for (int i = 0; i < N; ++i) p[(1 + i) * stride] = x.d[i];
  }
};
