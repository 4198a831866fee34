/* Oracle (test infrastructure only): restatement of the reference's level-1
 * column-pair kernels, one precision per inclusion (see vjo_svd.c for the
 * macro parameters).  Lane structure (VL accumulators, gcc's reduction tree)
 * is kept where it decides the rounding, so that every kernel reproduces the
 * reference's AVX-512 result bit for bit; lengths need not be multiples of
 * VL here (lanes past the end are simply not updated).
 */

/* The GPU's column norm, restated for comparison runs (vjo_norm_mode 1):
 * scale by 2^-k, k the exponent of max|x_i| over re and im, accumulate the
 * squares exactly enough (double-double for FP64, double for FP32 data),
 * take one corrected square root and return (e,f) = (k + e(r), f(r)).  The
 * result is the correctly rounded norm except within a few ulp^2 of a
 * rounding tie, so it matches the GPU whatever the summation order. */
static T SFX(norm_cr)(const vjo_fnat m, const T *const xr, const T *const xi, T *const e0, T *const f0)
{
  T mx = (T)0;
  for (vjo_fnat i = 0u; i < m; ++i) {
    mx = SFX(xmax)(mx, SFX(abs)(xr[i]));
    if (xi)
      mx = SFX(xmax)(mx, SFX(abs)(xi[i]));
  }
  if (!(mx > (T)0)) {
    *e0 = -(T)HUGE_VAL;
    *f0 = (T)1;
    return (T)0;
  }
  if (!(mx <= (T)TMAX)) {
    *e0 = (T)HUGE_VAL;
    *f0 = (T)1;
    return mx;
  }
  const int k = ilogb((double)mx);
  double s = 0.0, c = 0.0;
  for (vjo_fnat i = 0u; i < m; ++i) {
    for (int part = 0; part < (xi ? 2 : 1); ++part) {
      const double a = ldexp((double)(part ? xi[i] : xr[i]), -k);
      if (sizeof(T) == sizeof(double)) {
        const double p = a * a;
        const double e = fma(a, a, -p);
        double t;
        two_sum_d(s, p, &s, &t);
        c += (t + e);
      }
      else
        s += a * a;
    }
  }
  double hi, lo;
  two_sum_d(s, c, &hi, &lo);
  double r = sqrt(hi);
  if (r > 0.0) {
    const double d = fma(-r, r, hi) + lo;
    r = r + d / (2.0 * r);
  }
  const T rt = (T)r;
  const int q = ilogb((double)rt);
  *e0 = (T)(k + q);
  *f0 = (T)ldexp((double)rt, -q);
  return (T)ldexp((double)rt, k);
}

/* ?nrm2 with (e,f) split: restates dnormf_ (dnormf.c:5-16, the reference's
 * long-double norm) as the column norm; dnorm2_ (dnorm2.c:9-14) delegates
 * the same quantity to a BLAS ?nrm2 that is not part of the reference. */
T RN(norm2)(const vjo_fnat m, const T *const x, T *const e0, T *const f0)
{
  if (vjo_norm_mode)
    return SFX(norm_cr)(m, x, (const T*)NULL, e0, f0);
  long double sq = 0.0L;
  for (vjo_fnat i = 0u; i < m; ++i) {
    const long double l = x[i];
    sq += (l * l);
  }
  sq = sqrtl(sq);
  SFX(ldbl2ef)(sq, e0, f0);
  return (T)sq;
}

/* complex column: the same long-double accumulation over re and im
 * (znorm2.c:9-15 takes hypot of two BLAS norms; the value is the same
 * quantity, the rounding path is this oracle's own) */
T CN(norm2)(const vjo_fnat m, const T *const xr, const T *const xi, T *const e0, T *const f0)
{
  if (vjo_norm_mode)
    return SFX(norm_cr)(m, xr, xi, e0, f0);
  long double sq = 0.0L;
  for (vjo_fnat i = 0u; i < m; ++i) {
    const long double lr = xr[i];
    const long double li = xi[i];
    sq += (lr * lr);
    sq += (li * li);
  }
  sq = sqrtl(sq);
  SFX(ldbl2ef)(sq, e0, f0);
  return (T)sq;
}

/* The GPU's scaled dot product, restated for comparison runs (mode 1):
 * elements scaled by 2^-k with k the exponent of each column's max |.|,
 * products accumulated exactly enough (double-double), the sum rounded once
 * and rescaled by the exact power of two that turns the max-exponent
 * scaling into the reference's norm-exponent scaling (ddpscl.c:36-37), then
 * divided by fx fy as the reference does (ddpscl.c:73-74).  Complex:
 * x^H y.  The result is the reference's quantity correctly rounded (up to
 * ties), independent of the summation order. */
static void SFX(dot_cr)(const vjo_fnat m, const T *const xr, const T *const xi, const T *const yr, const T *const yi,
                        const T *const e, const T *const f, T *const zr, T *const zi)
{
  *zr = (T)0;
  if (zi)
    *zi = (T)0;
  if (!(e[0u] >= -(T)TMAX) || !(e[1u] >= -(T)TMAX))
    return;
  T mx = (T)0, my = (T)0;
  for (vjo_fnat i = 0u; i < m; ++i) {
    mx = SFX(xmax)(mx, SFX(abs)(xr[i]));
    my = SFX(xmax)(my, SFX(abs)(yr[i]));
    if (xi) {
      mx = SFX(xmax)(mx, SFX(abs)(xi[i]));
      my = SFX(xmax)(my, SFX(abs)(yi[i]));
    }
  }
  const int kx = ilogb((double)mx), ky = ilogb((double)my);
  double sr = 0.0, cr = 0.0, si = 0.0, ci = 0.0;
#define VJO_ACC(S, C, P)                        \
  {                                             \
    const double p_ = (P);                      \
    double t_;                                  \
    two_sum_d(S, p_, &S, &t_);                  \
    C += t_;                                    \
  }
  for (vjo_fnat i = 0u; i < m; ++i) {
    const double a = ldexp((double)xr[i], -kx), b = ldexp((double)yr[i], -ky);
    /* exact products as a + e with e = fma(.,.,-p); for FP32 data the
     * product is exact in double already */
    double p = a * b;
    VJO_ACC(sr, cr, p);
    cr += fma(a, b, -p);
    if (xi) {
      const double c = ldexp((double)xi[i], -kx), d = ldexp((double)yi[i], -ky);
      p = c * d;
      VJO_ACC(sr, cr, p);
      cr += fma(c, d, -p);
      p = a * d;
      VJO_ACC(si, ci, p);
      ci += fma(a, d, -p);
      p = -(c * b);
      VJO_ACC(si, ci, p);
      ci -= fma(c, b, -(c * b));
    }
  }
#undef VJO_ACC
  const int shift = (kx + ky) - (int)(e[0u] + e[1u]);
  const T de = (f[0u] * f[1u]);
  *zr = (T)ldexp((double)(T)(sr + cr), shift) / de;
  if (zi)
    *zi = (T)ldexp((double)(T)(si + ci), shift) / de;
}

/* ddpscl_ (ddpscl.c:7-75) with USE_2SUM */
T RN(dpscl)(const vjo_fnat m, const T *const x, const T *const y, const T *const e, const T *const f)
{
  if (vjo_norm_mode) {
    T z = (T)0;
    SFX(dot_cr)(m, x, (const T*)NULL, y, (const T*)NULL, e, f, &z, (T*)NULL);
    return z;
  }
  const T ex = e[0u];
  if (!(ex >= -(T)TMAX))
    return (T)0;
  const T ey = e[1u];
  if (!(ey >= -(T)TMAX))
    return (T)0;
  const T xe = -ex, ye = -ey;
  T spd[VL], sdn[VL], tpd[VL], tdn[VL];
  for (unsigned k = 0u; k < VL; ++k)
    spd[k] = sdn[k] = tpd[k] = tdn[k] = (T)0;
  for (vjo_fnat i = 0u; i < m; i += VL) {
    for (unsigned k = 0u; k < VL; ++k) {
      if (i + k >= m)
        break;
      const T xi = SFX(scalef)(x[i + k], xe);
      const T yi = SFX(scalef)(y[i + k], ye);
      const T pd = SFX(mul_rd)(xi, yi);
      const T dn = FMA(xi, yi, -pd);
      SFX(two_sum)(spd[k], (pd + tpd[k]), spd + k, tpd + k);
      SFX(two_sum)(sdn[k], (dn + tdn[k]), sdn + k, tdn + k);
    }
  }
  const T nu = (RED(spd) + (RED(tpd) + (RED(sdn) + RED(tdn))));
  const T de = (f[0u] * f[1u]);
  return (nu / de);
}

/* zdpscl_ (zdpscl.c:21-116) with USE_2SUM: x^H y / (|x| |y|) */
void CN(dpscl)(const vjo_fnat m, const T *const xr, const T *const xi, const T *const yr, const T *const yi, const T *const e, const T *const f, T *const zr, T *const zi)
{
  *zr = *zi = (T)0;
  if (vjo_norm_mode) {
    SFX(dot_cr)(m, xr, xi, yr, yi, e, f, zr, zi);
    return;
  }
  const T ex = e[0u];
  if (!(ex >= -(T)TMAX))
    return;
  const T ey = e[1u];
  if (!(ey >= -(T)TMAX))
    return;
  const T xe = -ex, ye = -ey;
  T sr[VL], tr[VL], sr_[VL], tr_[VL], si[VL], ti[VL], si_[VL], ti_[VL];
  for (unsigned k = 0u; k < VL; ++k)
    sr[k] = tr[k] = sr_[k] = tr_[k] = si[k] = ti[k] = si_[k] = ti_[k] = (T)0;
#define VJO_MUL2SUM(X, Y, S, TT, S_, T_) {            \
    const T xy = SFX(mul_rd)((X), (Y));               \
    const T xy_ = FMA((X), (Y), -xy);                 \
    SFX(two_sum)(S[k], (xy + TT[k]), S + k, TT + k);  \
    SFX(two_sum)(S_[k], (xy_ + T_[k]), S_ + k, T_ + k); }
  for (vjo_fnat i = 0u; i < m; i += VL) {
    for (unsigned k = 0u; k < VL; ++k) {
      if (i + k >= m)
        break;
      const T xri = SFX(scalef)(xr[i + k], xe);
      const T yri = SFX(scalef)(yr[i + k], ye);
      const T yii = SFX(scalef)(yi[i + k], ye);
      T xii = SFX(scalef)(xi[i + k], xe);
      VJO_MUL2SUM(xri, yri, sr, tr, sr_, tr_);
      VJO_MUL2SUM(xii, yii, sr, tr, sr_, tr_);
      xii = SFX(neg)(xii);
      VJO_MUL2SUM(xri, yii, si, ti, si_, ti_);
      VJO_MUL2SUM(xii, yri, si, ti, si_, ti_);
    }
  }
#undef VJO_MUL2SUM
  const T nur = (RED(sr) + (RED(tr) + (RED(sr_) + RED(tr_))));
  const T nui = (RED(si) + (RED(ti) + (RED(si_) + RED(ti_))));
  const T de = (f[0u] * f[1u]);
  *zr = nur / de;
  *zi = nui / de;
}

/* djrot_ (djrot.c:5-107) / djrotf_ (djrotf.c:5-100); track selects the
 * max-|element| return of the tracking form.  c != 1 scales by c, or
 * divides by it (the secant) in the USE_SECANTS form (VDMULDIV, vecdef.h:97-101). */
static T RN(jrot_any)(const int track, const vjo_fint n, T *const x, T *const y, const T c, const T at)
{
  if (!n)
    return (T)0;
  T mx = (T)0;
  if (n < 0) {
    const vjo_fnat n_ = (vjo_fnat)-n;
    if (at == (T)0) {
      if (c == (T)1) {
        for (vjo_fnat i = 0u; i < n_; ++i) { const T t = x[i]; x[i] = y[i]; y[i] = t; }
        return (T)0;
      }
      return (T)-4;
    }
    for (vjo_fnat i = 0u; i < n_; ++i) {
      const T x_ = x[i], y_ = y[i];
      T x_r = FMA(-x_, at, y_);
      T y_r = FMA(y_, at, x_);
      if (c != (T)1) { x_r = VJO_CSCALE(x_r, c); y_r = VJO_CSCALE(y_r, c); }
      mx = SFX(xmax)(mx, SFX(xmax)(SFX(abs)(x_r), SFX(abs)(y_r)));
      x[i] = x_r;
      y[i] = y_r;
    }
  }
  else {
    const vjo_fnat n_ = (vjo_fnat)n;
    if (at == (T)0)
      return ((c == (T)1) ? (T)0 : (T)-4);
    for (vjo_fnat i = 0u; i < n_; ++i) {
      const T x_ = x[i], y_ = y[i];
      T x_r = FMA(y_, at, x_);
      T y_r = FMA(-x_, at, y_);
      if (c != (T)1) { x_r = VJO_CSCALE(x_r, c); y_r = VJO_CSCALE(y_r, c); }
      mx = SFX(xmax)(mx, SFX(xmax)(SFX(abs)(x_r), SFX(abs)(y_r)));
      x[i] = x_r;
      y[i] = y_r;
    }
  }
  return (track ? mx : (T)0);
}
T RN(jrot)(const vjo_fint n, T *const x, T *const y, const T c, const T at) { return RN(jrot_any)(1, n, x, y, c, at); }
T RN(jrotf)(const vjo_fint n, T *const x, T *const y, const T c, const T at) { return RN(jrot_any)(0, n, x, y, c, at); }

/* zjrot_ (zjrot.c:6-186) / zjrotf_ (zjrotf.c:6-151): x' = (x + y tau) c,
 * y' = (y - x conj(tau)) c with tau = cat + i sat; VZFMA order (vecdef.h:74-77) */
static T CN(jrot_any)(const int track, const vjo_fint n, T *const xr, T *const xi, T *const yr, T *const yi, const T c, const T cat, const T sat)
{
  if (!n)
    return (T)0;
  if (sat == (T)0) {
    if (track) {
      const T r_r = RN(jrot)(n, xr, yr, c, cat);
      if (!(r_r >= (T)0))
        return (T)-10;
      const T r_i = RN(jrot)(n, xi, yi, c, cat);
      if (!(r_i >= (T)0))
        return (T)-11;
      return FMAX(r_r, r_i);
    }
    if (RN(jrotf)(n, xr, yr, c, cat) != (T)0)
      return (T)-10;
    if (RN(jrotf)(n, xi, yi, c, cat) != (T)0)
      return (T)-11;
    return (T)0;
  }
  const T cna = -cat;
  T mx = (T)0;
  const vjo_fnat n_ = (vjo_fnat)((n < 0) ? -n : n);
  for (vjo_fnat i = 0u; i < n_; ++i) {
    const T x_r = xr[i], x_i = xi[i], y_r = yr[i], y_i = yi[i];
    /* first VZFMA(_r,_i, y_r,y_i, cta,sta, x_r,x_i) */
    T ar = FMA(y_r, cat, FMA(-y_i, sat, x_r));
    T ai = FMA(y_r, sat, FMA(y_i, cat, x_i));
    if (c != (T)1) { ar = VJO_CSCALE(ar, c); ai = VJO_CSCALE(ai, c); }
    /* second VZFMA(_r,_i, x_r,x_i, cna,sta, y_r,y_i) */
    T br = FMA(x_r, cna, FMA(-x_i, sat, y_r));
    T bi = FMA(x_r, sat, FMA(x_i, cna, y_i));
    if (c != (T)1) { br = VJO_CSCALE(br, c); bi = VJO_CSCALE(bi, c); }
    mx = SFX(xmax)(mx, SFX(xmax)(SFX(abs)(ar), SFX(abs)(ai)));
    mx = SFX(xmax)(mx, SFX(xmax)(SFX(abs)(br), SFX(abs)(bi)));
    if (n < 0) { yr[i] = ar; yi[i] = ai; xr[i] = br; xi[i] = bi; }
    else { xr[i] = ar; xi[i] = ai; yr[i] = br; yi[i] = bi; }
  }
  return (track ? mx : (T)0);
}
T CN(jrot)(const vjo_fint n, T *const xr, T *const xi, T *const yr, T *const yi, const T c, const T cat, const T sat) { return CN(jrot_any)(1, n, xr, xi, yr, yi, c, cat, sat); }
T CN(jrotf)(const vjo_fint n, T *const xr, T *const xi, T *const yr, T *const yi, const T c, const T cat, const T sat) { return CN(jrot_any)(0, n, xr, xi, yr, yi, c, cat, sat); }

/* dgsscl_ (dgsscl.c:5-85): one-sided Gram-Schmidt for disparate norms */
T RN(gsscl)(const vjo_fint m, const T t, T *const x, T *const y, const T *const e, const T *const f)
{
  if (!m)
    return (T)0;
  const T ex = e[0u];
  if (!(ex >= -(T)TMAX))
    return (T)0;
  const T ey = e[1u];
  if (!(ey >= -(T)TMAX))
    return (T)0;
  const T fx = f[0u], fy = f[1u];
  const T x_e = -ex, y_e = -ey;
  T M = (T)0;
  if (m < 0) {
    const T tf = -t * (fx / fy);
    const vjo_fnat _m = (vjo_fnat)-m;
    for (vjo_fnat i = 0u; i < _m; ++i) {
      T xi = SFX(scalef)(x[i], x_e);
      T yi = y[i];
      x[i] = yi;
      yi = SFX(scalef)(yi, y_e);
      xi = SFX(scalef)(FMA(tf, yi, xi), ex);
      y[i] = xi;
      M = SFX(xmax)(M, SFX(abs)(xi));
    }
  }
  else {
    const T tf = -t * (fy / fx);
    const vjo_fnat _m = (vjo_fnat)m;
    for (vjo_fnat i = 0u; i < _m; ++i) {
      const T xi = SFX(scalef)(x[i], x_e);
      T yi = SFX(scalef)(y[i], y_e);
      yi = SFX(scalef)(FMA(tf, xi, yi), ey);
      y[i] = yi;
      M = SFX(xmax)(M, SFX(abs)(yi));
    }
  }
  return M;
}

/* zgsscl_ (zgsscl.c:5-119) */
T CN(gsscl)(const vjo_fint m, const T tr, const T ti, T *const xr, T *const xi, T *const yr, T *const yi, const T *const e, const T *const f)
{
  if (!m)
    return (T)0;
  const T ex = e[0u];
  if (!(ex >= -(T)TMAX))
    return (T)0;
  const T ey = e[1u];
  if (!(ey >= -(T)TMAX))
    return (T)0;
  const T fx = f[0u], fy = f[1u];
  const T x_e = -ex, y_e = -ey;
  T M = (T)0;
  if (m < 0) {
    const T fx_y = (fx / fy);
    const T tfr = -tr * fx_y;
    const T tfi = -ti * fx_y;
    const vjo_fnat _m = (vjo_fnat)-m;
    for (vjo_fnat i = 0u; i < _m; ++i) {
      const T yri = yr[i], yii = yi[i], xri = xr[i], xii = xi[i];
      xr[i] = yri;
      xi[i] = yii;
      const T yri_ = SFX(scalef)(yri, y_e);
      const T yii_ = SFX(scalef)(yii, y_e);
      const T xri_ = SFX(scalef)(xri, x_e);
      const T xii_ = SFX(scalef)(xii, x_e);
      /* VZFMA(yri,yii, tfr,tfi, yri_,yii_, xri_,xii_) */
      const T nr = FMA(tfr, yri_, FMA(-tfi, yii_, xri_));
      const T ni = FMA(tfr, yii_, FMA(tfi, yri_, xii_));
      const T outr = SFX(scalef)(nr, ex);
      const T outi = SFX(scalef)(ni, ex);
      yr[i] = outr;
      yi[i] = outi;
      M = SFX(xmax)(M, SFX(abs)(outr));
      M = SFX(xmax)(M, SFX(abs)(outi));
    }
  }
  else {
    const T fy_x = (fy / fx);
    const T tfr = -tr * fy_x;
    const T tfi = ti * fy_x;
    const vjo_fnat _m = (vjo_fnat)m;
    for (vjo_fnat i = 0u; i < _m; ++i) {
      const T xri_ = SFX(scalef)(xr[i], x_e);
      const T xii_ = SFX(scalef)(xi[i], x_e);
      const T yri_ = SFX(scalef)(yr[i], y_e);
      const T yii_ = SFX(scalef)(yi[i], y_e);
      /* VZFMA(xri,xii, tfr,tfi, xri_,xii_, yri_,yii_) */
      const T nr = FMA(tfr, xri_, FMA(-tfi, xii_, yri_));
      const T ni = FMA(tfr, xii_, FMA(tfi, xri_, yii_));
      const T outr = SFX(scalef)(nr, ey);
      const T outi = SFX(scalef)(ni, ey);
      yr[i] = outr;
      yi[i] = outi;
      M = SFX(xmax)(M, SFX(abs)(outr));
      M = SFX(xmax)(M, SFX(abs)(outi));
    }
  }
  return M;
}

/* dswp_ (dswp.c:3-26) */
void RN(swp)(const vjo_fnat n, T *const x, T *const y)
{
  for (vjo_fnat i = 0u; i < n; ++i) { const T t = x[i]; x[i] = y[i]; y[i] = t; }
}

/* dscale_ (dscale.c:5-33, dzscal.h:8-25): A *= 2^e exactly (scalef) */
void RN(scale)(const vjo_fnat m, const vjo_fnat n, T *const A, const vjo_fnat ldA, const vjo_fint e)
{
  if (!e)
    return;
  const T e_ = (T)e;
  for (vjo_fnat j = 0u; j < n; ++j)
    for (vjo_fnat i = 0u; i < m; ++i)
      A[j * ldA + i] = SFX(scalef)(A[j * ldA + i], e_);
}

/* dnormx_ (dnormx.c:5-36, dznrmx.h:10-27): max |a_ij|, NaN counted as inf */
T RN(normx)(const vjo_fnat m, const vjo_fnat n, const T *const A, const vjo_fnat ldA)
{
  T x = (T)0;
  for (vjo_fnat j = 0u; j < n; ++j)
    for (vjo_fnat i = 0u; i < m; ++i)
      x = SFX(xmax)(SFX(xmin)(SFX(abs)(A[j * ldA + i]), (T)HUGE_VAL), x);
  return x;
}
