/* Oracle (test infrastructure only): scalar restatements of the AVX-512 lane
 * primitives the reference kernels are built from.  Only tests/, the smoke
 * check and the benchmark's CPU-baseline leg may call into oracle/.
 *
 * Every primitive reproduces the x86 instruction semantics bit for bit,
 * including the NaN and signed-zero behaviour that the reference's
 * branch-free code relies on:
 *   xmin/xmax  = vminpd/vmaxpd   (a<b ? a : b ; a>b ? a : b; second operand on NaN)
 *   getexp     = vgetexppd       (floor(log2|x|); -inf for 0, +inf for inf)
 *   getmant    = vgetmantpd with _MM_MANT_NORM_1_2, _MM_MANT_SIGN_zero (1 for 0 and inf)
 *   scalef     = vscalefpd       (x * 2^floor(y), correctly rounded)
 *   mul_rd     = vmulpd with _MM_FROUND_TO_NEG_INF (ddpscl.c:52, zdpscl.c:10)
 *   two_sum    = TwoSum          (d2sum.h:8-14)
 *   red8/red16 = gcc's _mm512_reduce_add_pd/_ps tree (hi half + lo half, twice)
 */
#ifndef VJO_PRIM_H
#define VJO_PRIM_H

#include <float.h>
#include <math.h>
#include <stdint.h>
#include <string.h>

static inline uint64_t vjo_bits_d(double x) { uint64_t u; memcpy(&u, &x, 8); return u; }
static inline double vjo_dbl(uint64_t u) { double x; memcpy(&x, &u, 8); return x; }
static inline uint32_t vjo_bits_f(float x) { uint32_t u; memcpy(&u, &x, 4); return u; }
static inline float vjo_flt(uint32_t u) { float x; memcpy(&x, &u, 4); return x; }

/* ---------------- double ---------------- */
static inline double xmin_d(double a, double b) { return (a < b) ? a : b; }
static inline double xmax_d(double a, double b) { return (a > b) ? a : b; }
static inline double abs_d(double x) { return vjo_dbl(vjo_bits_d(x) & ~(UINT64_C(1) << 63)); }
static inline double sgn_d(double x) { return vjo_dbl(vjo_bits_d(x) & (UINT64_C(1) << 63)); }
static inline double or_d(double a, double b) { return vjo_dbl(vjo_bits_d(a) | vjo_bits_d(b)); }
static inline double xor_d(double a, double b) { return vjo_dbl(vjo_bits_d(a) ^ vjo_bits_d(b)); }
static inline double neg_d(double x) { return xor_d(x, -0.0); }

static inline double getexp_d(double x)
{
  if (isnan(x)) return x + x;
  if (x == 0.0) return -HUGE_VAL;
  if (isinf(x)) return HUGE_VAL;
  return (double)ilogb(x);
}

static inline double getmant_d(double x)
{
  if (isnan(x)) return x + x;
  if ((x == 0.0) || isinf(x)) return 1.0;
  return ldexp(abs_d(x), -ilogb(x));
}

static inline double scalef_d(double x, double y)
{
  if (isnan(x)) return (y == HUGE_VAL) ? HUGE_VAL : ((y == -HUGE_VAL) ? 0.0 : (x + x));
  if (isnan(y)) return y + y;
  if (isinf(y)) {
    if (isinf(x)) return (y > 0.0) ? x : (x - x);
    if (x == 0.0) return (y > 0.0) ? (y - y) : x;
    return (y > 0.0) ? copysign(HUGE_VAL, x) : copysign(0.0, x);
  }
  double f = floor(y);
  if (f > 4096.0) f = 4096.0;
  if (f < -4096.0) f = -4096.0;
  return ldexp(x, (int)f);
}

static inline double mul_rd_d(double x, double y)
{
  double p = x * y;
  if (!isfinite(p)) return p;
  const double r = fma(x, y, -p);
  if ((r < 0.0) || ((r == 0.0) && signbit(r)))
    p = nextafter(p, -HUGE_VAL);
  return p;
}

static inline void two_sum_d(double a, double b, double *s, double *t)
{
  const double s_ = a + b;
  double a_ = s_ - b;
  double b_ = s_ - a_;
  a_ = a - a_;
  b_ = b - b_;
  *s = s_;
  *t = a_ + b_;
}

static inline double red8_d(const double v[8])
{
  double r4[4], r2[2];
  for (int i = 0; i < 4; ++i) r4[i] = v[i + 4] + v[i];
  for (int i = 0; i < 2; ++i) r2[i] = r4[i + 2] + r4[i];
  return r2[0] + r2[1];
}

/* (e,f) lexicographic comparisons: VDEFLT/VDEFLE (defops.h:31-37) */
static inline int eflt_d(double e0, double e1, double f0, double f1) { return (e0 < e1) || ((e0 == e1) && (f0 < f1)); }
static inline int efle_d(double e0, double e1, double f0, double f1) { return (e0 < e1) || ((e0 == e1) && (f0 <= f1)); }
/* efcmp (defops.h:46-56) */
static inline int efcmp_d(double e0, double f0, double e1, double f1)
{
  if (e0 > e1) return 1;
  if (e0 < e1) return -1;
  if (f0 > f1) return 1;
  if (f0 < f1) return -1;
  return 0;
}
/* dbl2ef (dznrm2.h:4-22) */
static inline void dbl2ef_d(double d, double *e, double *f)
{
  if (isfinite(d)) {
    int de = 0;
    const double df = frexp(d, &de);
    if (df == 0.0) { *e = -HUGE_VAL; *f = copysign(1.0, df); }
    else { *e = (double)(de - 1); *f = scalbn(df, 1); }
  }
  else { *e = HUGE_VAL; *f = copysign((isinf(d) ? 1.0 : 0.0), d); }
}
/* ldbl2ef (dznrmf.h:4-22) */
static inline void ldbl2ef_d(long double l, double *e, double *f)
{
  if (isfinite(l)) {
    int le = 0;
    const double df = (double)frexpl(l, &le);
    if (df == 0.0) { *e = -HUGE_VAL; *f = copysign(1.0, df); }
    else { *e = (double)(le - 1); *f = scalbn(df, 1); }
  }
  else { *e = HUGE_VAL; *f = (double)copysignl((isinf(l) ? 1.0L : 0.0L), l); }
}

/* ---------------- float ---------------- */
static inline float xmin_f(float a, float b) { return (a < b) ? a : b; }
static inline float xmax_f(float a, float b) { return (a > b) ? a : b; }
static inline float abs_f(float x) { return vjo_flt(vjo_bits_f(x) & ~(UINT32_C(1) << 31)); }
static inline float sgn_f(float x) { return vjo_flt(vjo_bits_f(x) & (UINT32_C(1) << 31)); }
static inline float or_f(float a, float b) { return vjo_flt(vjo_bits_f(a) | vjo_bits_f(b)); }
static inline float xor_f(float a, float b) { return vjo_flt(vjo_bits_f(a) ^ vjo_bits_f(b)); }
static inline float neg_f(float x) { return xor_f(x, -0.0f); }

static inline float getexp_f(float x)
{
  if (isnan(x)) return x + x;
  if (x == 0.0f) return -HUGE_VALF;
  if (isinf(x)) return HUGE_VALF;
  return (float)ilogbf(x);
}

static inline float getmant_f(float x)
{
  if (isnan(x)) return x + x;
  if ((x == 0.0f) || isinf(x)) return 1.0f;
  return ldexpf(abs_f(x), -ilogbf(x));
}

static inline float scalef_f(float x, float y)
{
  if (isnan(x)) return (y == HUGE_VALF) ? HUGE_VALF : ((y == -HUGE_VALF) ? 0.0f : (x + x));
  if (isnan(y)) return y + y;
  if (isinf(y)) {
    if (isinf(x)) return (y > 0.0f) ? x : (x - x);
    if (x == 0.0f) return (y > 0.0f) ? (y - y) : x;
    return (y > 0.0f) ? copysignf(HUGE_VALF, x) : copysignf(0.0f, x);
  }
  float f = floorf(y);
  if (f > 512.0f) f = 512.0f;
  if (f < -512.0f) f = -512.0f;
  return ldexpf(x, (int)f);
}

static inline float mul_rd_f(float x, float y)
{
  float p = x * y;
  if (!isfinite(p)) return p;
  const float r = fmaf(x, y, -p);
  if ((r < 0.0f) || ((r == 0.0f) && signbit(r)))
    p = nextafterf(p, -HUGE_VALF);
  return p;
}

static inline void two_sum_f(float a, float b, float *s, float *t)
{
  const float s_ = a + b;
  float a_ = s_ - b;
  float b_ = s_ - a_;
  a_ = a - a_;
  b_ = b - b_;
  *s = s_;
  *t = a_ + b_;
}

static inline float red16_f(const float v[16])
{
  float r8[8], r4[4];
  for (int i = 0; i < 8; ++i) r8[i] = v[i + 8] + v[i];
  for (int i = 0; i < 4; ++i) r4[i] = r8[i + 4] + r8[i];
  const float c0 = r4[0] + r4[2];
  const float c1 = r4[1] + r4[3];
  return c0 + c1;
}

static inline int eflt_f(float e0, float e1, float f0, float f1) { return (e0 < e1) || ((e0 == e1) && (f0 < f1)); }
static inline int efle_f(float e0, float e1, float f0, float f1) { return (e0 < e1) || ((e0 == e1) && (f0 <= f1)); }
static inline int efcmp_f(float e0, float f0, float e1, float f1)
{
  if (e0 > e1) return 1;
  if (e0 < e1) return -1;
  if (f0 > f1) return 1;
  if (f0 < f1) return -1;
  return 0;
}
/* flt2ef (scnrm2.h:4-22) */
static inline void flt2ef_f(float d, float *e, float *f)
{
  if (isfinite(d)) {
    int de = 0;
    const float df = frexpf(d, &de);
    if (df == 0.0f) { *e = -HUGE_VALF; *f = copysignf(1.0f, df); }
    else { *e = (float)(de - 1); *f = scalbnf(df, 1); }
  }
  else { *e = HUGE_VALF; *f = copysignf((isinf(d) ? 1.0f : 0.0f), d); }
}
static inline void dbl2ef_f(float d, float *e, float *f) { flt2ef_f(d, e, f); }
static inline void ldbl2ef_f(long double l, float *e, float *f)
{
  if (isfinite(l)) {
    int le = 0;
    const float df = (float)frexpl(l, &le);
    if (df == 0.0f) { *e = -HUGE_VALF; *f = copysignf(1.0f, df); }
    else { *e = (float)(le - 1); *f = scalbnf(df, 1); }
  }
  else { *e = HUGE_VALF; *f = (float)copysignl((isinf(l) ? 1.0L : 0.0L), l); }
}

#endif /* !VJO_PRIM_H */
