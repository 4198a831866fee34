/* Oracle (test infrastructure only): restatement of the one-sided Jacobi SVD
 * drivers.  Included four times by vjo_svd.c (CPLX 0/1 x double/float).
 *   dvjsvd_ dvjsvd.c:40-638   zvjsvd_ zvjsvd.c:40-694
 *   svjsvd_ svjsvd.c:40-638   cvjsvd_ cvjsvd.c:40-694
 * The phases of every step are kept in the reference's order (norms with the
 * overflow loop, scaled dots, convergence test and Gram prescaling, 2x2 EVD,
 * action codes, apply), so that error returns and the rescaling path behave
 * as in the reference.  Two deliberate differences, neither visible on a
 * successful run with a step table whose rows are permutations:
 *   - the "norm is stale" flags live in their own array instead of aliasing
 *     the w1/w2 workspace (dvjsvd.c:169-180, 402);
 *   - the workspace is allocated here; work[] and iwork[1..] are untouched.
 * strict=1 applies the reference's argument checks (multiples of VL,
 * 64-byte alignment, MXCSR); strict=0 accepts any even n and any m >= n.
 * The norm, dot and apply loops over the pairs of a step run on a thread
 * team with max/min reductions, as the reference's do (dvjsvd.c:244, 303,
 * 463): the pairs of a step touch disjoint columns, so a successful run is
 * bit-identical to a serial one.  Inside an enclosing parallel region (the
 * batch helper) they run serially. */
#ifndef VJO_PAR
#define VJO_PAR(M, N) (((M) * (N)) >= 32768u)
#endif

#if CPLX
#define DRVNAME CN(vjsvd_core)
#define DRVPUB CN(vjsvd)
#define ROT_EXP ((int)ROT_EXP_C)
#define L_MG 92
#define L_MV 104
#define L_DOT_R 345
#define L_DOT_I 347
#define L_APP_G 630
#define L_APP_V 638
#define RC_FPENV (-18)
#define RC_WORK (-16)
#define RC_VZERO (-17)
#else
#define DRVNAME RN(vjsvd_core)
#define DRVPUB RN(vjsvd)
#define ROT_EXP ((int)ROT_EXP_R)
#define L_MG 80
#define L_MV 92
#define L_DOT_R 320
#define L_APP_G 580
#define L_APP_V 588
#define RC_FPENV (-14)
#define RC_WORK (-12)
#define RC_VZERO (-13)
#endif

vjo_fint DRVNAME(const int strict, const vjo_fnat *const m, const vjo_fnat *const n,
                 T *const Gr, const vjo_fnat *const ldGr, T *const Gi, const vjo_fnat *const ldGi,
                 T *const Vr, const vjo_fnat *const ldVr, T *const Vi, const vjo_fnat *const ldVi,
                 T *const eS, T *const fS, const unsigned *const js, const unsigned *const stp,
                 const unsigned *const swp, T *const work, unsigned *const iwork, vjo_stats *const st_out)
{
  const vjo_fnat M = *m, N = *n, n_2 = (N >> 1u);
  if (strict && vjo_fpenv_bad())
    return RC_FPENV;
  if (!N)
    return 0;
  if (M < N)
    return -1;
  if (strict && (M % VL))
    return -1;
  if (N & 1u)
    return -2;
  if (strict && (n_2 % VL))
    return -2;
#if CPLX
  if (strict && VJO_MISALIGNED(Gr)) return -3;
  if (*ldGr < M) return -4;
  if (strict && (*ldGr % VL)) return -4;
  if (strict && VJO_MISALIGNED(Gi)) return -5;
  if (*ldGi < M) return -6;
  if (strict && (*ldGi % VL)) return -6;
  if (strict && VJO_MISALIGNED(Vr)) return -7;
  if (*ldVr < N) return -8;
  if (strict && (*ldVr % VL)) return -8;
  if (strict && VJO_MISALIGNED(Vi)) return -9;
  if (*ldVi < N) return -10;
  if (strict && (*ldVi % VL)) return -10;
#else
  (void)Gi; (void)ldGi; (void)Vi; (void)ldVi;
  if (strict && VJO_MISALIGNED(Gr)) return -3;
  if (*ldGr < M) return -4;
  if (strict && (*ldGr % VL)) return -4;
  if (strict && VJO_MISALIGNED(Vr)) return -5;
  if (*ldVr < N) return -6;
  if (strict && (*ldVr % VL)) return -6;
#endif
  if (strict && VJO_MISALIGNED(work))
    return RC_WORK;

  const vjo_fnat lG = *ldGr, lV = *ldVr;
#if CPLX
  const vjo_fnat lGi = *ldGi, lVi = *ldVi;
#define GR(j) (Gr + (size_t)(j) * lG)
#define GI(j) (Gi + (size_t)(j) * lGi)
#define VR(j) (Vr + (size_t)(j) * lV)
#define VI(j) (Vi + (size_t)(j) * lVi)
#define NORMX_G() FMAX(RN(normx)(M, N, Gr, lG), RN(normx)(M, N, Gi, lGi))
#define NORMX_V() FMAX(RN(normx)(N, N, Vr, lV), RN(normx)(N, N, Vi, lVi))
#define SCALE_G(s) { RN(scale)(M, N, Gr, lG, (s)); RN(scale)(M, N, Gi, lGi, (s)); }
#define SCALE_V(s) { RN(scale)(N, N, Vr, lV, (s)); RN(scale)(N, N, Vi, lVi, (s)); }
#else
#define GR(j) (Gr + (size_t)(j) * lG)
#define VR(j) (Vr + (size_t)(j) * lV)
#define NORMX_G() RN(normx)(M, N, Gr, lG)
#define NORMX_V() RN(normx)(N, N, Vr, lV)
#define SCALE_G(s) RN(scale)(M, N, Gr, lG, (s))
#define SCALE_V(s) RN(scale)(N, N, Vr, lV, (s))
#endif

  T MG = NORMX_G();
  if (!(MG <= (T)TMAX))
    return -L_MG;
  if (COPYSIGN((T)1, MG) == (T)-1)
    return -(L_MG + 2);

  T MV = (T)1;
  if (*iwork) {
    MV = NORMX_V();
    if (!(MV <= (T)TMAX))
      return -L_MV;
    if (MV <= (T)0)
      return RC_VZERO;
  }
  else {
    for (vjo_fnat j = 0u; j < N; ++j) {
      for (vjo_fnat i = 0u; i < N; ++i) {
        VR(j)[i] = (T)0;
#if CPLX
        VI(j)[i] = (T)0;
#endif
      }
      VR(j)[j] = (T)1;
    }
  }

  for (vjo_fnat j = 0u; j < N; ++j) {
    eS[j] = -(T)HUGE_VAL;
    fS[j] = (T)1;
  }
  if (st_out) {
    st_out->sweeps = 0u;
    st_out->sT = 0;
    st_out->sV = 0;
    st_out->rescales = 0u;
  }

  if (MG == (T)0)
    return 0;
#if CPLX
  const T M_m = ((T)TMAX / ((T)(M << 2u) * (T)M_SQRT2));
#else
  const T M_m = ((T)TMAX / (T)(M << 1u));
#endif
  T es = (T)0, fs = (T)0;
  SFX(dbl2ef)(M_m, &es, &fs);
  const int NRM_EXP = (int)es;
  SFX(dbl2ef)(MG, &es, &fs);
  int eM = (int)es;
#if CPLX
  int sR = ROT_EXP - eM - 1;
#else
  int sR = ROT_EXP - eM;
#endif
  vjo_fint sN = NRM_EXP - eM - 1;
  if (sN) {
    SCALE_G(sN);
    MG = SCALBN(MG, (int)sN);
  }
  int sT = (int)sN;

  if (*iwork) {
    SFX(dbl2ef)(MV, &es, &fs);
    eM = (int)es;
#if CPLX
    sR = ROT_EXP - eM - 1;
#else
    sR = ROT_EXP - eM;
#endif
    sN = sR;
    if (sN) {
      SCALE_V(sN);
      MV = SCALBN(MV, sR);
    }
  }
  else
    sR = 0;
  int sV = sR;

  /* workspace (dvjsvd.c:161-172, zvjsvd.c:176-190) */
  T *const ws = (T*)calloc(16u * (size_t)(n_2 + 1u) + 2u * (size_t)N, sizeof(T));
  if (!ws)
    return -__LINE__;
  T *const a11 = ws;
  T *const a22 = a11 + n_2;
  T *const a21 = a22 + n_2;
  T *const a21i = a21 + n_2;
  T *const cc = a21i + n_2;
  T *const at = cc + n_2;
  T *const sat = at + n_2;
  T *const l1 = sat + n_2;
  T *const l2 = l1 + n_2;
  T *const code = l2 + n_2;
  T *const stale = code + n_2;

  if (*swp)
    for (vjo_fnat i = 0u; i < N; ++i)
      stale[i] = (T)1;

  /* see LAPACK's xGESVJ (dvjsvd.c:182-184) */
  const T tol = SQRT((T)M) * SCALBN((T)TEPS, -1);
  const T gst = SCALBN(tol, (int)FIN_EXP);
  unsigned sw = 0u;
  vjo_fint rc = 0;

  while (sw < *swp) {
    size_t swt = 0u;
    for (unsigned st = 0u; st < *stp; ++st) {
      /* rescale G according to MG if necessary (dvjsvd.c:198-222) */
      SFX(dbl2ef)(MG, &es, &fs);
      eM = (int)es;
#if CPLX
      sR = ROT_EXP - eM - 1;
#else
      sR = ROT_EXP - eM;
#endif
      sN = NRM_EXP - eM - 1;
      if (sR < 0) {
        SCALE_G(sN);
        sR = (int)sN;
        MG = SCALBN(MG, sR);
        sT += sR;
        for (vjo_fnat i = 0u; i < N; ++i)
          stale[i] = (T)1;
        if (st_out) ++st_out->rescales;
      }
      /* rescale V according to MV if necessary (dvjsvd.c:223-233) */
      SFX(dbl2ef)(MV, &es, &fs);
      eM = (int)es;
#if CPLX
      sR = ROT_EXP - eM - 1;
#else
      sR = ROT_EXP - eM;
#endif
      sN = sR;
      if (sR < 0) {
        SCALE_V(sN);
        MV = SCALBN(MV, sR);
        sV += sR;
      }
      const unsigned *const r = js + st * (size_t)N;
      /* norms, overflow-aware (dvjsvd.c:234-296) */
      T nMG = (T)-0.0;
      int overflow = 0;
      do {
        nMG = (T)0;
        /* one thread per pair block, as the reference's loop (dvjsvd.c:244) */
#pragma omp parallel for schedule(static) reduction(max:nMG) if(VJO_PAR(M, N))
        for (vjo_fnat pq = 0u; pq < N; pq += 2u) {
          const vjo_fnat k = (pq >> 1u);
          if (!(nMG <= (T)TMAX)) {
            a11[k] = (T)NAN;
            a22[k] = (T)NAN;
            continue;
          }
          const size_t p = r[pq], q = r[pq + 1u];
          if (stale[p] == (T)1) {
#if CPLX
            a11[k] = CN(norm2)(M, GR(p), GI(p), eS + p, fS + p);
#else
            a11[k] = RN(norm2)(M, GR(p), eS + p, fS + p);
#endif
            nMG = FMAX(nMG, FMIN(a11[k], (T)HUGE_VAL));
            if (!(nMG <= (T)TMAX)) {
              a22[k] = (T)NAN;
              continue;
            }
          }
          else
            a11[k] = SFX(scalef)(fS[p], eS[p]);
          if (stale[q] == (T)1) {
#if CPLX
            a22[k] = CN(norm2)(M, GR(q), GI(q), eS + q, fS + q);
#else
            a22[k] = RN(norm2)(M, GR(q), eS + q, fS + q);
#endif
            nMG = FMAX(nMG, FMIN(a22[k], (T)HUGE_VAL));
          }
          else
            a22[k] = SFX(scalef)(fS[q], eS[q]);
        }
        if ((overflow = !(nMG <= (T)TMAX))) {
          SCALE_G(sN);
          sR = (int)sN;
          MG = SCALBN(MG, sR);
          sT += sR;
          for (vjo_fnat i = 0u; i < N; ++i)
            stale[i] = (T)1;
          if (st_out) ++st_out->rescales;
        }
      } while (overflow);
      /* scaled dot-products (dvjsvd.c:297-331) */
      nMG = (T)0;
#pragma omp parallel for schedule(static) reduction(min:nMG) if(VJO_PAR(M, N)) /* dvjsvd.c:303 */
      for (vjo_fnat pq = 0u; pq < N; pq += 2u) {
        const vjo_fnat k = (pq >> 1u);
        if (!(nMG >= (T)0)) {
          a21[k] = (T)NAN;
          continue;
        }
        const size_t p = r[pq], q = r[pq + 1u];
        const T e2[2] = { eS[q], eS[p] };
        const T f2[2] = { fS[q], fS[p] };
#if CPLX
        CN(dpscl)(M, GR(q), GI(q), GR(p), GI(p), e2, f2, a21 + k, a21i + k);
        nMG = FMIN(nMG, (isfinite(a21[k]) ? (T)0 : (T)-L_DOT_R));
        nMG = FMIN(nMG, (isfinite(a21i[k]) ? (T)0 : (T)-L_DOT_I));
#else
        a21[k] = RN(dpscl)(M, GR(q), GR(p), e2, f2);
        nMG = FMIN(nMG, (isfinite(a21[k]) ? (T)0 : (T)-L_DOT_R));
#endif
      }
      if (!(nMG >= (T)0)) {
        rc = (vjo_fint)nMG;
        goto done;
      }
      /* convergence test, Gram prescaling and the 2x2 EVD
       * (dvjsvd.c:349-417); the reference evaluates a whole block of VL
       * pairs when any of them transforms, which only affects lanes whose
       * results are never used. */
      size_t stt = 0u;
      for (vjo_fnat k = 0u; k < n_2; ++k) {
        const size_t p = r[2u * k], q = r[2u * k + 1u];
#if CPLX
        T aabs;
        {
          const T x_ = SFX(abs)(a21[k]), y_ = SFX(abs)(a21i[k]);
          const T m_ = SFX(xmin)(x_, y_), M_ = SFX(xmax)(x_, y_);
          aabs = m_ / M_;
          aabs = SFX(xmax)(aabs, (T)0);
          aabs = FMA(aabs, aabs, (T)1);
          aabs = SQRT(aabs);
          aabs = aabs * M_;
        }
#else
        const T aabs = SFX(abs)(a21[k]);
#endif
        const int trans = (tol <= aabs);
        if (trans) {
          ++stt;
          const T e1 = eS[p], e2 = eS[q], f1 = fS[p], f2 = fS[q];
          const int ngs = SFX(eflt)(e1, e2, f1, f2);
          T maf = ngs ? f1 : f2;
          T mae = ngs ? e1 : e2;
          const T Maf = ngs ? f2 : f1;
          const T Mae = ngs ? e2 : e1;
          maf = gst * maf;
          mae = mae + SFX(getexp)(maf);
          maf = SFX(getmant)(maf);
          const int cgs = SFX(eflt)(mae, Mae, maf, Maf);
          const int gsp = (!ngs && cgs), gsn = (ngs && cgs);
          T f12 = f1 / f2, e12 = e1 - e2, f21 = f2 / f1, e21 = e2 - e1;
          e12 = e12 + SFX(getexp)(f12);
          f12 = SFX(getmant)(f12);
          e21 = e21 + SFX(getexp)(f21);
          f21 = SFX(getmant)(f21);
          const int c12 = SFX(efle)(e12, e21, f12, f21);
          const T E = c12 ? e21 : e12;
          const T d = SFX(xmin)((T)FIN_EXP - E, (T)0);
          e12 = e12 + d;
          e21 = e21 + d;
          const T w0 = SFX(scalef)(f12, e12);
          const T w1 = SFX(scalef)(f21, e21);
          const T w2 = SFX(scalef)(a21[k], d);
          unsigned P = 0u;
#if CPLX
          const T w3 = SFX(scalef)(a21i[k], d);
          SFX(cjac2)(vjo_secants ? VJO_SEC : VJO_TAN, w0, w1, w2, w3, cc + k, at + k, sat + k, l1 + k, l2 + k, &P);
#else
          SFX(rjac2)(vjo_secants ? VJO_SEC : VJO_TAN, w0, w1, w2, cc + k, at + k, l1 + k, l2 + k, &P);
#endif
          /* action code (dvjsvd.c:438-447) */
          code[k] = gsp ? (T)3 : (gsn ? (T)-3 : (P ? (T)-2 : (T)2));
        }
        else if (SFX(efcmp)(eS[p], fS[p], eS[q], fS[q]) < 0) {
          T t = eS[p]; eS[p] = eS[q]; eS[q] = t;
          t = fS[p]; fS[p] = fS[q]; fS[q] = t;
          code[k] = (T)-1;
        }
        else
          code[k] = (T)1;
      }
      swt += stt;
      /* apply (dvjsvd.c:460-590) */
      nMG = (T)0;
      T nMV = (T)-0.0;
#pragma omp parallel for schedule(static) reduction(max:nMG, nMV) if(VJO_PAR(M, N)) /* dvjsvd.c:463 */
      for (vjo_fnat k = 0u; k < n_2; ++k) {
        const size_t p = r[2u * k], q = r[2u * k + 1u];
        stale[q] = stale[p] = (T)0;
        if (!(nMG <= (T)TMAX)) {
          nMG = (T)HUGE_VAL;
          continue;
        }
        if (!(nMV <= (T)TMAX)) {
          nMV = (T)HUGE_VAL;
          continue;
        }
        const T w = code[k];
        if (w == (T)-3) {
          const T e2[2] = { eS[p], eS[q] };
          const T f2[2] = { fS[p], fS[q] };
#if CPLX
          const T tG = CN(gsscl)(-(vjo_fint)M, a21[k], a21i[k], GR(p), GI(p), GR(q), GI(q), e2, f2);
#else
          const T tG = RN(gsscl)(-(vjo_fint)M, a21[k], GR(p), GR(q), e2, f2);
#endif
          if (!isfinite(tG)) {
            nMG = (T)HUGE_VAL;
            continue;
          }
          nMG = FMAX(nMG, tG);
          RN(swp)(N, VR(p), VR(q));
#if CPLX
          RN(swp)(N, VI(p), VI(q));
#endif
          nMV = FMAX(nMV, (T)0);
          T t = eS[p]; eS[p] = eS[q]; eS[q] = t;
          t = fS[p]; fS[p] = fS[q]; fS[q] = t;
          stale[q] = (T)1;
          continue;
        }
        else if ((w == (T)-2) || (w == (T)2)) {
          const vjo_fint _m = ((w < (T)0) ? -(vjo_fint)M : (vjo_fint)M);
          const vjo_fint _n = ((w < (T)0) ? -(vjo_fint)N : (vjo_fint)N);
#if CPLX
          const T tG = CN(jrot)(_m, GR(p), GI(p), GR(q), GI(q), cc[k], at[k], sat[k]);
#else
          const T tG = RN(jrot)(_m, GR(p), GR(q), cc[k], at[k]);
#endif
          if (!isfinite(tG)) {
            nMG = (T)HUGE_VAL;
            continue;
          }
          nMG = FMAX(nMG, tG);
#if CPLX
          const T tV = (*iwork ? CN(jrot) : CN(jrotf))(_n, VR(p), VI(p), VR(q), VI(q), cc[k], at[k], sat[k]);
#else
          const T tV = (*iwork ? RN(jrot) : RN(jrotf))(_n, VR(p), VR(q), cc[k], at[k]);
#endif
          if (!isfinite(tV)) {
            nMV = (T)HUGE_VAL;
            continue;
          }
          nMV = FMAX(nMV, tV);
        }
        else if (w == (T)-1) {
          RN(swp)(M, GR(p), GR(q));
#if CPLX
          RN(swp)(M, GI(p), GI(q));
#endif
          nMG = FMAX(nMG, (T)0);
          RN(swp)(N, VR(p), VR(q));
#if CPLX
          RN(swp)(N, VI(p), VI(q));
#endif
          nMV = FMAX(nMV, (T)0);
          continue;
        }
        else if (w == (T)1) {
          nMG = FMAX(nMG, (T)0);
          nMV = FMAX(nMV, (T)0);
          continue;
        }
        else if (w == (T)3) {
          const T e2[2] = { eS[p], eS[q] };
          const T f2[2] = { fS[p], fS[q] };
#if CPLX
          const T tG = CN(gsscl)((vjo_fint)M, a21[k], a21i[k], GR(p), GI(p), GR(q), GI(q), e2, f2);
#else
          const T tG = RN(gsscl)((vjo_fint)M, a21[k], GR(p), GR(q), e2, f2);
#endif
          if (!isfinite(tG)) {
            nMG = (T)HUGE_VAL;
            continue;
          }
          nMG = FMAX(nMG, tG);
          nMV = FMAX(nMV, (T)0);
          stale[q] = (T)1;
          continue;
        }
        else {
          nMG = (T)HUGE_VAL;
          nMV = (T)HUGE_VAL;
          continue;
        }
        stale[q] = stale[p] = (T)1;
      }
      if (!(nMG <= (T)TMAX)) {
        rc = -L_APP_G;
        goto done;
      }
      MG = FMAX(MG, nMG);
      if (!(nMV <= (T)TMAX)) {
        rc = -L_APP_V;
        goto done;
      }
      MV = FMAX(MV, nMV);
    }
    if (st_out && (sw < VJO_MAX_TRACE))
      st_out->swt[sw] = (uint64_t)swt;
    if (!swt)
      break;
    ++sw;
  }

  /* epilogue (dvjsvd.c:604-631): G <- U, sigma = fS * 2^(eS - sT) */
  if (sw < *swp) {
    for (vjo_fnat j = 0u; j < N; ++j) {
      const T f_ = fS[j], s_ = -eS[j];
      for (vjo_fnat i = 0u; i < M; ++i) {
        GR(j)[i] = SFX(scalef)(GR(j)[i] / f_, s_);
#if CPLX
        GI(j)[i] = SFX(scalef)(GI(j)[i] / f_, s_);
#endif
      }
      eS[j] -= (T)sT;
    }
    if (sV) {
      const T s_ = (T)-sV;
      for (vjo_fnat j = 0u; j < N; ++j)
        for (vjo_fnat i = 0u; i < N; ++i) {
          VR(j)[i] = SFX(scalef)(VR(j)[i], s_);
#if CPLX
          VI(j)[i] = SFX(scalef)(VI(j)[i], s_);
#endif
        }
    }
  }
  rc = (vjo_fint)sw;
  if (st_out) {
    st_out->sweeps = sw;
    st_out->sT = sT;
    st_out->sV = sV;
  }
 done:
  free(ws);
  return rc;
#undef GR
#undef VR
#ifdef GI
#undef GI
#undef VI
#endif
#undef NORMX_G
#undef NORMX_V
#undef SCALE_G
#undef SCALE_V
}

#if CPLX
vjo_fint DRVPUB(const vjo_fnat *m, const vjo_fnat *n, T *Gr, const vjo_fnat *ldGr, T *Gi, const vjo_fnat *ldGi, T *Vr, const vjo_fnat *ldVr, T *Vi, const vjo_fnat *ldVi, T *eS, T *fS, const unsigned *js, const unsigned *stp, const unsigned *swp, T *work, unsigned *iwork)
{
  return DRVNAME(1, m, n, Gr, ldGr, Gi, ldGi, Vr, ldVr, Vi, ldVi, eS, fS, js, stp, swp, work, iwork, (vjo_stats*)NULL);
}
#else
vjo_fint DRVPUB(const vjo_fnat *m, const vjo_fnat *n, T *G, const vjo_fnat *ldG, T *V, const vjo_fnat *ldV, T *eS, T *fS, const unsigned *js, const unsigned *stp, const unsigned *swp, T *work, unsigned *iwork)
{
  return DRVNAME(1, m, n, G, ldG, (T*)NULL, ldG, V, ldV, (T*)NULL, ldV, eS, fS, js, stp, swp, work, iwork, (vjo_stats*)NULL);
}
#endif

#undef DRVNAME
#undef DRVPUB
#undef ROT_EXP
#undef L_MG
#undef L_MV
#undef L_DOT_R
#ifdef L_DOT_I
#undef L_DOT_I
#endif
#undef L_APP_G
#undef L_APP_V
#undef RC_FPENV
#undef RC_WORK
#undef RC_VZERO
