/* Oracle (test infrastructure only): per-lane restatement of the batched
 * 2x2 Hermitian Jacobi kernels.  Included once per precision by vjo_jac2.c
 * with these macros defined:
 *   T        element type (double / float)
 *   SFX(x)   appends the primitive suffix (_d / _f)
 *   RN(x)    builds the real public name  (vjo_dbjac2 / vjo_sbjac2)
 *   CN(x)    builds the complex public name (vjo_zbjac2 / vjo_cbjac2)
 *   VL       AVX-512 lanes of T (8 / 16): the bit layout of p[]
 *   BIG_EXP  DBL_BIG_EXP 1020 (dzjac2.h:10) / FLT_BIG_EXP 124 (scjac2.h:10)
 *   SQRT_HUGE DBL_SQRT_HUGE (dzjac2.h:16) / FLT_SQRT_HUGE (scjac2.h:16)
 *   TMAX, TMIN_SUB  DBL_MAX / DBL_TRUE_MIN (or FLT_)
 *   FMA      fma / fmaf ; SQRT sqrt / sqrtf
 *
 * Forms (what the second and third outputs hold):
 *   VJO_TAN  c = cos, at = tan       D8JAC2_LOOP  (dbjac2.c:17-44), Z8JAC2_LOOP (zbjac2.c:18-58)
 *   VJO_SIN  c = cos, at = sin       D8JACL_LOOP  (dbjac2.c:51-78), Z8JACL_LOOP (zbjac2.c:65-105)
 *   VJO_SEC  c = sec, at = tan       D8JACI_LOOP  (dbjac2.c:148-174), Z8JACI_LOOP (zbjac2.c:179-218)
 */

static inline void SFX(rjac2)(const int form, const T a11, const T a22, const T a21, T *const c, T *const at, T *const l1, T *const l2, unsigned *const P)
{
  T a1 = a11, a2 = a22, ar = a21;
  const T e1 = (T)BIG_EXP - SFX(getexp)(a1);
  const T e2 = (T)BIG_EXP - SFX(getexp)(a2);
  const T er = (T)BIG_EXP - SFX(getexp)(ar);
  T es = SFX(xmin)(SFX(xmin)(e1, e2), SFX(xmin)(er, (T)TMAX));
  ar = SFX(scalef)(ar, es);
  a1 = SFX(scalef)(a1, es);
  a2 = SFX(scalef)(a2, es);
  const T aa = SFX(abs)(ar);
  const T as = SFX(sgn)(ar);
  es = SFX(neg)(es);
  const T ab = SFX(scalef)(aa, (T)1);
  const T ad = a1 - a2;
  const T t2 = SFX(or)(SFX(xmin)(SFX(xmax)(ab / SFX(abs)(ad), (T)0), (T)SQRT_HUGE), SFX(sgn)(ad));
  const T t1 = t2 / ((T)1 + SQRT(FMA(t2, t2, (T)1)));
  const T s2 = FMA(t1, t1, (T)1);
  const T s1 = SQRT(s2);
  switch (form) {
  case VJO_SIN:
    *at = SFX(xor)(t1, as) / s1;
    *c = (T)1 / s1;
    break;
  case VJO_SEC:
    *at = SFX(xor)(t1, as);
    *c = s1;
    break;
  default:
    *at = SFX(xor)(t1, as);
    *c = (T)1 / s1;
    break;
  }
  const T L1 = FMA(t1, FMA(a2, t1, ab), a1) / s2;
  const T L2 = FMA(t1, FMA(a1, t1, -ab), a2) / s2;
  *l1 = SFX(scalef)(L1, es);
  *l2 = SFX(scalef)(L2, es);
  *P = (L1 < L2);
}

static inline void SFX(cjac2)(const int form, const T a11, const T a22, const T a21r, const T a21i, T *const c, T *const cat, T *const sat, T *const l1, T *const l2, unsigned *const P)
{
  T a1 = a11, a2 = a22, ar = a21r, ai = a21i;
  const T e1 = (T)BIG_EXP - SFX(getexp)(a1);
  const T e2 = (T)BIG_EXP - SFX(getexp)(a2);
  const T er = (T)BIG_EXP - SFX(getexp)(ar);
  const T ei = (T)BIG_EXP - SFX(getexp)(ai);
  T es = SFX(xmin)((T)TMAX, SFX(xmin)(SFX(xmin)(e1, e2), SFX(xmin)(er, ei)));
  ar = SFX(scalef)(ar, es);
  ai = SFX(scalef)(ai, es);
  a1 = SFX(scalef)(a1, es);
  a2 = SFX(scalef)(a2, es);
  T ar_ = SFX(abs)(ar);
  T ai_ = SFX(abs)(ai);
  const T as = SFX(sgn)(ar);
  const T am = SFX(xmin)(ar_, ai_);
  const T aM = SFX(xmax)(ar_, ai_);
  es = SFX(neg)(es);
  T aa = SFX(xmax)(am / aM, (T)0);
  aa = SQRT(FMA(aa, aa, (T)1)) * aM;
  ar_ = SFX(or)(SFX(xmin)(ar_ / aa, (T)1), as);
  ai_ = ai / SFX(xmax)(aa, (T)TMIN_SUB);
  const T ab = SFX(scalef)(aa, (T)1);
  const T ad = a1 - a2;
  const T t2 = SFX(or)(SFX(xmin)(SFX(xmax)(ab / SFX(abs)(ad), (T)0), (T)SQRT_HUGE), SFX(sgn)(ad));
  const T t1 = t2 / ((T)1 + SQRT(FMA(t2, t2, (T)1)));
  const T s2 = FMA(t1, t1, (T)1);
  const T s1 = SQRT(s2);
  switch (form) {
  case VJO_SIN:
    *c = (T)1 / s1;
    *cat = (ar_ * t1) / s1;
    *sat = (ai_ * t1) / s1;
    break;
  case VJO_SEC:
    *c = s1;
    *cat = ar_ * t1;
    *sat = ai_ * t1;
    break;
  default:
    *c = (T)1 / s1;
    *cat = ar_ * t1;
    *sat = ai_ * t1;
    break;
  }
  const T L1 = FMA(t1, FMA(a2, t1, ab), a1) / s2;
  const T L2 = FMA(t1, FMA(a1, t1, -ab), a2) / s2;
  *l1 = SFX(scalef)(L1, es);
  *l2 = SFX(scalef)(L2, es);
  *P = (L1 < L2);
}

/* Public batch entry: n<0 sines, n>0 tangents (dbjac2.c:104-141,
 * zbjac2.c:108-170).  p[i/VL] receives the L1<L2 bits of lanes i..i+VL-1.
 * Returns 0 (the reference's serial build) or -1 when |n| is not a multiple
 * of VL (the reference's debug-build argument check, dbjac2.c:86-87). */
vjo_fint RN(bjac2)(const vjo_fint *const n, const T *const a11, const T *const a22, const T *const a21, T *const c, T *const at, T *const l1, T *const l2, unsigned *const p)
{
  const int form = (*n < 0) ? VJO_SIN : VJO_TAN;
  const vjo_fnat nn = (vjo_fnat)((*n < 0) ? -*n : *n);
  if (nn % VL)
    return -1;
  for (vjo_fnat i = 0u; i < nn; i += VL) {
    unsigned w = 0u;
    for (unsigned k = 0u; k < VL; ++k) {
      unsigned P = 0u;
      SFX(rjac2)(form, a11[i + k], a22[i + k], a21[i + k], c + i + k, at + i + k, l1 + i + k, l2 + i + k, &P);
      w |= (P << k);
    }
    p[i / VL] = w;
  }
  return 0;
}

vjo_fint CN(bjac2)(const vjo_fint *const n, const T *const a11, const T *const a22, const T *const a21r, const T *const a21i, T *const c, T *const cat, T *const sat, T *const l1, T *const l2, unsigned *const p)
{
  const int form = (*n < 0) ? VJO_SIN : VJO_TAN;
  const vjo_fnat nn = (vjo_fnat)((*n < 0) ? -*n : *n);
  if (nn % VL)
    return -1;
  for (vjo_fnat i = 0u; i < nn; i += VL) {
    unsigned w = 0u;
    for (unsigned k = 0u; k < VL; ++k) {
      unsigned P = 0u;
      SFX(cjac2)(form, a11[i + k], a22[i + k], a21r[i + k], a21i[i + k], c + i + k, cat + i + k, sat + i + k, l1 + i + k, l2 + i + k, &P);
      w |= (P << k);
    }
    p[i / VL] = w;
  }
  return 0;
}

/* Internal batch entry used by the SVD drivers: n<0 secants, n>=0 cosines;
 * a block of VL lanes is processed only when its low p-bits are non-zero,
 * and those low bits are replaced by L1<L2 (dbjac2.c:178-251,
 * zbjac2.c:222-299). */
vjo_fint RN(bjac2i)(const vjo_fint *const n, const T *const a11, const T *const a22, const T *const a21, T *const c, T *const at, T *const l1, T *const l2, unsigned *const p)
{
  const int form = (*n < 0) ? VJO_SEC : VJO_TAN;
  const vjo_fnat nn = (vjo_fnat)((*n < 0) ? -*n : *n);
  const unsigned lo = (unsigned)((VL == 8u) ? 0xFFu : 0xFFFFu);
  if (nn % VL)
    return -1;
  for (vjo_fnat i = 0u; i < nn; i += VL) {
    const vjo_fnat j = i / VL;
    if (!(p[j] & lo))
      continue;
    unsigned w = 0u;
    for (unsigned k = 0u; k < VL; ++k) {
      unsigned P = 0u;
      SFX(rjac2)(form, a11[i + k], a22[i + k], a21[i + k], c + i + k, at + i + k, l1 + i + k, l2 + i + k, &P);
      w |= (P << k);
    }
    p[j] = ((p[j] & ~lo) | w);
  }
  return 0;
}

vjo_fint CN(bjac2i)(const vjo_fint *const n, const T *const a11, const T *const a22, const T *const a21r, const T *const a21i, T *const c, T *const cat, T *const sat, T *const l1, T *const l2, unsigned *const p)
{
  const int form = (*n < 0) ? VJO_SEC : VJO_TAN;
  const vjo_fnat nn = (vjo_fnat)((*n < 0) ? -*n : *n);
  const unsigned lo = (unsigned)((VL == 8u) ? 0xFFu : 0xFFFFu);
  if (nn % VL)
    return -1;
  for (vjo_fnat i = 0u; i < nn; i += VL) {
    const vjo_fnat j = i / VL;
    if (!(p[j] & lo))
      continue;
    unsigned w = 0u;
    for (unsigned k = 0u; k < VL; ++k) {
      unsigned P = 0u;
      SFX(cjac2)(form, a11[i + k], a22[i + k], a21r[i + k], a21i[i + k], c + i + k, cat + i + k, sat + i + k, l1 + i + k, l2 + i + k, &P);
      w |= (P << k);
    }
    p[j] = ((p[j] & ~lo) | w);
  }
  return 0;
}
