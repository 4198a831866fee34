// Device-side numeric primitives for the CDNA4 (gfx950) kernels.
//
// The reference is written against AVX-512 and leans on a handful of x86
// instruction semantics (vminpd/vmaxpd NaN rules, vgetexppd, vgetmantpd,
// vscalefpd).  These helpers give the same results on the GPU, per lane, so
// that the 2x2 EVD and the column-pair updates reproduce the reference's
// rounding exactly; they are plain FP64/FP32 VALU code, no MFMA.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace vj {

template <typename T> struct Lim;
template <> struct Lim<double> {
  typedef unsigned long long U;
  static constexpr double MAX = 1.7976931348623157e308;
  static constexpr double TRUE_MIN = 4.9406564584124654e-324;
  static constexpr double EPS = 2.220446049250313e-16;
  static constexpr double BIG_EXP = 1020.0;                      // dzjac2.h:10
  static constexpr double SQRT_HUGE = 1.34078079299425956E+154;  // dzjac2.h:16
  static constexpr double FIN_EXP = 1023.0;                      // defops.h:43
  static constexpr int ROT_EXP_R = 1022;                         // dvjsvd.c:23
  static constexpr int ROT_EXP_C = 1021;                         // zvjsvd.c:23
  static constexpr int VL = 8;                                   // AVX-512 lanes (p[] bit layout)
  static constexpr int CLAMP = 4096;
};
template <> struct Lim<float> {
  typedef unsigned U;
  static constexpr float MAX = 3.40282347e+38f;
  static constexpr float TRUE_MIN = 1.40129846e-45f;
  static constexpr float EPS = 1.1920929e-07f;
  static constexpr float BIG_EXP = 124.0f;                       // scjac2.h:10
  static constexpr float SQRT_HUGE = 1.844674297e+19f;           // scjac2.h:16
  static constexpr float FIN_EXP = 127.0f;
  static constexpr int ROT_EXP_R = 126;                          // svjsvd.c:23
  static constexpr int ROT_EXP_C = 125;                          // cvjsvd.c:23
  static constexpr int VL = 16;
  static constexpr int CLAMP = 512;
};

__device__ __forceinline__ unsigned long long bits(double x) { return __double_as_longlong(x); }
__device__ __forceinline__ double from_bits(unsigned long long u) { return __longlong_as_double((long long)u); }
__device__ __forceinline__ unsigned bits(float x) { return __float_as_uint(x); }
__device__ __forceinline__ float from_bits(unsigned u) { return __uint_as_float(u); }

template <typename T> __device__ __forceinline__ T sign_mask();
template <> __device__ __forceinline__ double sign_mask<double>() { return -0.0; }
template <> __device__ __forceinline__ float sign_mask<float>() { return -0.0f; }

// vminpd / vmaxpd: the second operand wins unless the first compares strictly
template <typename T> __device__ __forceinline__ T xmin(T a, T b) { return (a < b) ? a : b; }
template <typename T> __device__ __forceinline__ T xmax(T a, T b) { return (a > b) ? a : b; }
template <typename T> __device__ __forceinline__ T vabs(T x) { return from_bits(bits(x) & ~bits(sign_mask<T>())); }
template <typename T> __device__ __forceinline__ T vsgn(T x) { return from_bits(bits(x) & bits(sign_mask<T>())); }
template <typename T> __device__ __forceinline__ T vor(T a, T b) { return from_bits(bits(a) | bits(b)); }
template <typename T> __device__ __forceinline__ T vxor(T a, T b) { return from_bits(bits(a) ^ bits(b)); }
template <typename T> __device__ __forceinline__ T vneg(T x) { return vxor(x, sign_mask<T>()); }

__device__ __forceinline__ int frexp_exp(double x) { return __builtin_amdgcn_frexp_exp(x); }
__device__ __forceinline__ int frexp_exp(float x) { return __builtin_amdgcn_frexp_expf(x); }
__device__ __forceinline__ double frexp_mant(double x) { return __builtin_amdgcn_frexp_mant(x); }
__device__ __forceinline__ float frexp_mant(float x) { return __builtin_amdgcn_frexp_mantf(x); }
__device__ __forceinline__ double ldexp_(double x, int e) { return __builtin_amdgcn_ldexp(x, e); }
__device__ __forceinline__ float ldexp_(float x, int e) { return __builtin_amdgcn_ldexpf(x, e); }
__device__ __forceinline__ double fma_(double a, double b, double c) { return __builtin_fma(a, b, c); }
__device__ __forceinline__ float fma_(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
__device__ __forceinline__ double sqrt_(double x) { return __builtin_sqrt(x); }
__device__ __forceinline__ float sqrt_(float x) { return __builtin_sqrtf(x); }
__device__ __forceinline__ double floor_(double x) { return __builtin_floor(x); }
__device__ __forceinline__ float floor_(float x) { return __builtin_floorf(x); }
template <typename T> __device__ __forceinline__ bool isnan_(T x) { return x != x; }
template <typename T> __device__ __forceinline__ bool isinf_(T x) { return vabs(x) == (T)__builtin_inf(); }
template <typename T> __device__ __forceinline__ bool isfinite_(T x) { return vabs(x) <= Lim<T>::MAX; }
template <typename T> __device__ __forceinline__ T inf_() { return (T)__builtin_inf(); }

// vgetexppd: floor(log2|x|), -inf for 0, +inf for inf, quiet NaN for NaN
template <typename T> __device__ __forceinline__ T getexp(T x)
{
  const T ax = vabs(x);
  T r = (T)(frexp_exp(x) - 1);
  r = (ax == (T)0) ? -inf_<T>() : r;
  r = (ax == inf_<T>()) ? inf_<T>() : r;
  return isnan_(x) ? (x + x) : r;
}

// vgetmantpd(_MM_MANT_NORM_1_2, _MM_MANT_SIGN_zero): |mantissa| in [1,2); 1 for 0 and inf
template <typename T> __device__ __forceinline__ T getmant(T x)
{
  const T ax = vabs(x);
  T r = vabs(frexp_mant(x)) * (T)2;
  r = ((ax == (T)0) || (ax == inf_<T>())) ? (T)1 : r;
  return isnan_(x) ? (x + x) : r;
}

// x * 2^floor(y) for finite y (exact up to the final rounding)
template <typename T> __device__ __forceinline__ T scalef_fin(T x, T y)
{
  T f = floor_(y);
  f = xmin(xmax(f, (T)-Lim<T>::CLAMP), (T)Lim<T>::CLAMP);
  return ldexp_(x, (int)f);
}

// vscalefpd including its special cases (matters only for -inf/+inf exponents)
template <typename T> __device__ __forceinline__ T scalef(T x, T y)
{
  if (isfinite_(y))
    return isnan_(x) ? (x + x) : scalef_fin(x, y);
  if (isnan_(y))
    return y + y;
  const bool up = (y > (T)0);
  if (isnan_(x))
    return up ? inf_<T>() : (T)0;
  if (isinf_(x))
    return up ? x : (x - x);
  if (x == (T)0)
    return up ? (y - y) : x;
  return up ? vor(inf_<T>(), vsgn(x)) : vsgn(x);
}

// efcmp / VDEFLT / VDEFLE on (exponent, fraction) pairs (defops.h:31-56)
template <typename T> __device__ __forceinline__ bool eflt(T e0, T e1, T f0, T f1) { return (e0 < e1) || ((e0 == e1) && (f0 < f1)); }
template <typename T> __device__ __forceinline__ bool efle(T e0, T e1, T f0, T f1) { return (e0 < e1) || ((e0 == e1) && (f0 <= f1)); }
template <typename T> __device__ __forceinline__ int efcmp(T e0, T f0, T e1, T f1)
{
  return (e0 > e1) ? 1 : ((e0 < e1) ? -1 : ((f0 > f1) ? 1 : ((f0 < f1) ? -1 : 0)));
}

// Knuth/Moller TwoSum (d2sum.h:8-14): s + t == a + b exactly
template <typename T> __device__ __forceinline__ void two_sum(T a, T b, T &s, T &t)
{
  s = a + b;
  T a_ = s - b;
  T b_ = s - a_;
  a_ = a - a_;
  b_ = b - b_;
  t = a_ + b_;
}

// Form of the 2x2 outputs.
enum Jac2Form { JAC2_TAN = 0, JAC2_SIN = 1, JAC2_SEC = 2 };

// BIG_EXP - getexp(x) of a finite x as an integer, CLAMP for a zero x.
// getexp maps zero to -inf, so the reference's scale exponent min(e..., MAX)
// is an integer of magnitude below CLAMP (at most BIG_EXP + 1074 in FP64,
// BIG_EXP + 149 in FP32) when an input is nonzero, and MAX when all are
// zero; scalef_fin clamps MAX to CLAMP.  The integer minimum with zeros at
// CLAMP is therefore the exponent scalef_fin applies, and ldexp by it (or by
// its negation) gives the same bits.
template <typename T> __device__ __forceinline__ int jac2_ei(T x)
{
  return (x == (T)0) ? Lim<T>::CLAMP : ((int)Lim<T>::BIG_EXP + 1 - frexp_exp(x));
}

// scalef(x, y) for y = -e, e an integer exponent from jac2_ei (|e| <= CLAMP)
template <typename T> __device__ __forceinline__ T jac2_unscale(T x, int e)
{
  return isnan_(x) ? (x + x) : ldexp_(x, -e);
}

// One real symmetric 2x2 problem [[a11, a21], [a21, a22]] per lane.
// Same operation sequence as D8JAC2_LOOP / D8JACL_LOOP / D8JACI_LOOP
// (dbjac2.c:17-44, 51-78, 148-174): exact 2^es prescaling, tan(2phi) clipped
// to sqrt(huge), tan(phi), cos, and eigenvalues unscaled by 2^-es.
// FIN: the caller guarantees finite inputs (the SVD's pair decisions), so
// the special-value path is not compiled.
template <typename T, int FORM, bool FIN = false>
__device__ __forceinline__ void rjac2(T a11, T a22, T a21, T &c, T &at, T &l1, T &l2, bool &P)
{
  // finite inputs (every lane but the special-value ones): integer exponents
  const bool fin = FIN || (isfinite_(a11) && isfinite_(a22) && isfinite_(a21));
  int ei = 0;
  T es = (T)0, ar, a1, a2;
  if (fin) {
    ei = xmin(xmin(jac2_ei(a11), jac2_ei(a22)), jac2_ei(a21));
    ar = ldexp_(a21, ei);
    a1 = ldexp_(a11, ei);
    a2 = ldexp_(a22, ei);
  }
  else {
    const T e1 = (T)Lim<T>::BIG_EXP - getexp(a11);
    const T e2 = (T)Lim<T>::BIG_EXP - getexp(a22);
    const T er = (T)Lim<T>::BIG_EXP - getexp(a21);
    es = xmin(xmin(e1, e2), xmin(er, (T)Lim<T>::MAX));
    ar = scalef(a21, es);
    a1 = scalef(a11, es);
    a2 = scalef(a22, es);
    es = vneg(es);
  }
  const T aa = vabs(ar);
  const T as = vsgn(ar);
  const T ab = FIN ? ldexp_(aa, 1) : scalef(aa, (T)1);
  const T ad = a1 - a2;
  const T t2 = vor(xmin(xmax(ab / vabs(ad), (T)0), (T)Lim<T>::SQRT_HUGE), vsgn(ad));
  const T t1 = t2 / ((T)1 + sqrt_(fma_(t2, t2, (T)1)));
  const T s2 = fma_(t1, t1, (T)1);
  const T s1 = sqrt_(s2);
  if (FORM == JAC2_SIN) {
    at = vxor(t1, as) / s1;
    c = (T)1 / s1;
  }
  else if (FORM == JAC2_SEC) {
    at = vxor(t1, as);
    c = s1;
  }
  else {
    at = vxor(t1, as);
    c = (T)1 / s1;
  }
  const T L1 = fma_(t1, fma_(a2, t1, ab), a1) / s2;
  const T L2 = fma_(t1, fma_(a1, t1, -ab), a2) / s2;
  if (fin) {
    l1 = jac2_unscale(L1, ei);
    l2 = jac2_unscale(L2, ei);
  }
  else {
    l1 = scalef(L1, es);
    l2 = scalef(L2, es);
  }
  P = (L1 < L2);
}

// One Hermitian 2x2 problem [[a11, conj(a21)], [a21, a22]] per lane
// (Z8JAC2_LOOP / Z8JACL_LOOP / Z8JACI_LOOP, zbjac2.c:18-58, 65-105, 179-218):
// a21 = |a21| e^{i alpha} via a safe hypot, then the real path; outputs
// cat = cos(alpha) t, sat = sin(alpha) t (t = tan or sin per FORM).
template <typename T, int FORM, bool FIN = false>
__device__ __forceinline__ void cjac2(T a11, T a22, T a21r, T a21i, T &c, T &cat, T &sat, T &l1, T &l2, bool &P)
{
  // finite inputs (every lane but the special-value ones): integer exponents
  const bool fin = FIN || (isfinite_(a11) && isfinite_(a22) && isfinite_(a21r) && isfinite_(a21i));
  int ei = 0;
  T es = (T)0, ar, ai, a1, a2;
  if (fin) {
    ei = xmin(xmin(jac2_ei(a11), jac2_ei(a22)), xmin(jac2_ei(a21r), jac2_ei(a21i)));
    ar = ldexp_(a21r, ei);
    ai = ldexp_(a21i, ei);
    a1 = ldexp_(a11, ei);
    a2 = ldexp_(a22, ei);
  }
  else {
    const T e1 = (T)Lim<T>::BIG_EXP - getexp(a11);
    const T e2 = (T)Lim<T>::BIG_EXP - getexp(a22);
    const T er = (T)Lim<T>::BIG_EXP - getexp(a21r);
    const T eim = (T)Lim<T>::BIG_EXP - getexp(a21i);
    es = xmin((T)Lim<T>::MAX, xmin(xmin(e1, e2), xmin(er, eim)));
    ar = scalef(a21r, es);
    ai = scalef(a21i, es);
    a1 = scalef(a11, es);
    a2 = scalef(a22, es);
    es = vneg(es);
  }
  T ar_ = vabs(ar);
  T ai_ = vabs(ai);
  const T as = vsgn(ar);
  const T am = xmin(ar_, ai_);
  const T aM = xmax(ar_, ai_);
  T aa = xmax(am / aM, (T)0);
  aa = sqrt_(fma_(aa, aa, (T)1)) * aM;
  ar_ = vor(xmin(ar_ / aa, (T)1), as);
  ai_ = ai / xmax(aa, (T)Lim<T>::TRUE_MIN);
  const T ab = FIN ? ldexp_(aa, 1) : scalef(aa, (T)1);
  const T ad = a1 - a2;
  const T t2 = vor(xmin(xmax(ab / vabs(ad), (T)0), (T)Lim<T>::SQRT_HUGE), vsgn(ad));
  const T t1 = t2 / ((T)1 + sqrt_(fma_(t2, t2, (T)1)));
  const T s2 = fma_(t1, t1, (T)1);
  const T s1 = sqrt_(s2);
  if (FORM == JAC2_SIN) {
    c = (T)1 / s1;
    cat = (ar_ * t1) / s1;
    sat = (ai_ * t1) / s1;
  }
  else if (FORM == JAC2_SEC) {
    c = s1;
    cat = ar_ * t1;
    sat = ai_ * t1;
  }
  else {
    c = (T)1 / s1;
    cat = ar_ * t1;
    sat = ai_ * t1;
  }
  const T L1 = fma_(t1, fma_(a2, t1, ab), a1) / s2;
  const T L2 = fma_(t1, fma_(a1, t1, -ab), a2) / s2;
  if (fin) {
    l1 = jac2_unscale(L1, ei);
    l2 = jac2_unscale(L2, ei);
  }
  else {
    l1 = scalef(L1, es);
    l2 = scalef(L2, es);
  }
  P = (L1 < L2);
}

} // namespace vj
