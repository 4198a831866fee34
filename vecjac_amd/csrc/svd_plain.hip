// Plain (one matrix) one-sided Jacobi SVD on gfx950: the device side of
// dvjsvd_/zvjsvd_/svjsvd_/cvjsvd_ (dvjsvd.c:40-638, zvjsvd.c:40-694).
//
// One kernel launch per step of the strategy table: workgroup k owns pair
// (js[st][2k], js[st][2k+1]) and runs the whole step on it (norms, dot,
// convergence test, 2x2 EVD, update of the G and V columns).  The control
// state (max-|.| bounds, power-of-two shifts, per-sweep transformation
// counter, convergence flag, error code) lives in device memory and is
// updated by the last workgroup of each step, so sweeps run without any
// host round trip; once the flag is set, later launches return at once.
#include "svd_common.hpp"

namespace vj {

// ------------------------------------------------------------ setup kernels
template <typename T>
__global__ __launch_bounds__(256) void k_maxabs(const int m, const int n, const T *__restrict__ A, const long long ldA,
                                                unsigned long long *__restrict__ out)
{
  // max |a_ij| with NaN counted as inf (dznrmx.h:10-27)
  T v = (T)0;
  const long long total = (long long)m * n;
  for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < total; t += (long long)gridDim.x * blockDim.x) {
    const long long j = t / m, i = t - j * m;
    v = xmax(xmin(vabs(A[j * ldA + i]), inf_<T>()), v);
  }
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1)
    v = xmax(v, __shfl_xor(v, o, 64));
  if ((threadIdx.x & 63) == 0)
    atomicMax(out, (unsigned long long)__double_as_longlong((double)v));
}

template <typename T, bool CPLX>
__global__ void k_setup(const int m, PlainState<T> *st, const int vpreset)
{
  // dvjsvd.c:78-156 (zvjsvd.c:90-171): input checks and initial scaling
  if (threadIdx.x || blockIdx.x)
    return;
  const int L_MG = CPLX ? 92 : 80, L_MV = CPLX ? 104 : 92;
  const T MG = (T)__longlong_as_double((long long)st->MG_u);
  st->MG0 = MG;
  st->nev = 0u;
  st->err_sw = st->err_st = 0u;
  const T MV = vpreset ? (T)__longlong_as_double((long long)st->MV_u) : (T)1;
  st->sT = st->sV = 0;
  st->sG_next = st->sV_next = 0;
  st->sG_init = st->sV_init = 0;
  st->sweeps = 0;
  st->swt = 0;
  st->blocks_done = 0u;
  st->zero = 0;
  st->status = 0;
  st->done = 0;
  st->ef_init = 0;
  if (!(MG <= Lim<T>::MAX)) { st->status = -L_MG; st->done = 1; return; }
  if (vpreset) {
    if (!(MV <= Lim<T>::MAX)) { st->status = -L_MV; st->done = 1; return; }
    if (MV <= (T)0) { st->status = CPLX ? -17 : -13; st->done = 1; return; }
  }
  st->ef_init = 1; // V (unless preset) and eS/fS are initialised from here on
  if (MG == (T)0) { st->zero = 1; st->done = 1; return; }
  const T M_m = CPLX ? (Lim<T>::MAX / ((T)(m << 2) * (T)1.41421356237309504880)) : (Lim<T>::MAX / (T)(m << 1));
  const int NRM_EXP = frexp_exp(M_m) - 1;
  const int eM = frexp_exp(MG) - 1;
  const int sN = NRM_EXP - eM - 1;
  st->eM0 = eM;
  st->sR0 = (CPLX ? Lim<T>::ROT_EXP_C - 1 : Lim<T>::ROT_EXP_R) - eM;
  st->NRM_EXP = NRM_EXP;
  st->MG = ldexp_(MG, sN);
  st->sG_init = sN;
  st->sT = sN;
  int sV = 0;
  if (vpreset) {
    const int eV = frexp_exp(MV) - 1;
    sV = (CPLX ? Lim<T>::ROT_EXP_C - 1 : Lim<T>::ROT_EXP_R) - eV;
    st->MV = ldexp_(MV, sV);
  }
  else
    st->MV = (T)1;
  st->sV_init = sV;
  st->sV = sV;
}

template <typename T>
__global__ __launch_bounds__(256) void k_init_cols(const int m, const int n, T *__restrict__ Gr, T *__restrict__ Gi,
                                                   const long long ldG, T *__restrict__ Vr, T *__restrict__ Vi,
                                                   const long long ldV, T *__restrict__ eS, T *__restrict__ fS,
                                                   const PlainState<T> *__restrict__ st, const int vpreset)
{
  // column j = blockIdx.x: initial power-of-two scaling of G (dscale_), of a
  // preset V, V := I otherwise (dvjsvd.c:97-156), eS = -inf, fS = 1
  if (!st->ef_init)
    return;
  const int j = blockIdx.x;
  const int sG = st->zero ? 0 : st->sG_init;
  const int sV = st->sV_init;
  if (sG) {
    for (int i = threadIdx.x; i < m; i += blockDim.x) {
      Gr[j * ldG + i] = ldexp_(Gr[j * ldG + i], sG);
      if (Gi)
        Gi[j * ldG + i] = ldexp_(Gi[j * ldG + i], sG);
    }
  }
  for (int i = threadIdx.x; i < n; i += blockDim.x) {
    if (vpreset) {
      if (sV) {
        Vr[j * ldV + i] = ldexp_(Vr[j * ldV + i], sV);
        if (Vi)
          Vi[j * ldV + i] = ldexp_(Vi[j * ldV + i], sV);
      }
    }
    else {
      Vr[j * ldV + i] = (i == j) ? (T)1 : (T)0;
      if (Vi)
        Vi[j * ldV + i] = (T)0;
    }
  }
  if (threadIdx.x == 0) {
    eS[j] = -inf_<T>();
    fS[j] = (T)1;
  }
}

// ------------------------------------------------------------- step kernel
// Rows lane + j B (j < E) of the pair's columns, zero past m, scaled by
// 2^sG when a rescaling is pending.
template <typename T, bool CPLX, int B, int E>
__device__ __forceinline__ void load_rows(const T *Gp, const T *Gq, const T *Gpi, const T *Gqi, const int m, const int j0,
                                          const int sG, const int lane, T (&xr)[E], T (&yr)[E], T (&xi)[E], T (&yi)[E])
{
#pragma unroll
  for (int j = 0; j < E; ++j) {
    const int i = lane + (j0 + j) * B;
    const bool in = (i < m);
    xr[j] = in ? Gp[i] : (T)0;
    yr[j] = in ? Gq[i] : (T)0;
    if constexpr (CPLX) {
      xi[j] = in ? Gpi[i] : (T)0;
      yi[j] = in ? Gqi[i] : (T)0;
    }
    else {
      xi[j] = yi[j] = (T)0;
    }
    if (sG) {
      xr[j] = ldexp_(xr[j], sG);
      yr[j] = ldexp_(yr[j], sG);
      if constexpr (CPLX) {
        xi[j] = ldexp_(xi[j], sG);
        yi[j] = ldexp_(yi[j], sG);
      }
    }
  }
}

template <typename T, bool CPLX, int B, int E>
__device__ __forceinline__ void store_rows(T *Gp, T *Gq, T *Gpi, T *Gqi, const int m, const int j0, const int lane,
                                           const T (&xr)[E], const T (&yr)[E], const T (&xi)[E], const T (&yi)[E])
{
#pragma unroll
  for (int j = 0; j < E; ++j) {
    const int i = lane + (j0 + j) * B;
    if (i < m) {
      Gp[i] = xr[j];
      Gq[i] = yr[j];
      if constexpr (CPLX) {
        Gpi[i] = xi[j];
        Gqi[i] = yi[j];
      }
    }
  }
}

// One step on pair blockIdx.x.  STREAM = false: the pair is held in
// registers, E rows per lane.  STREAM = true (columns too long for the
// registers): the pair is read three times in chunks of E rows per lane -
// maxima, sums, update - with the same per-lane row order, hence the same
// bits as the register form with the same group size.
template <typename T, bool CPLX, int B, int E, bool STREAM = false>
__global__ __launch_bounds__(B) void k_step(const int m, const int n, T *__restrict__ Gr, T *__restrict__ Gi,
                                             const long long ldG, T *__restrict__ Vr, T *__restrict__ Vi,
                                             const long long ldV, T *__restrict__ eS, T *__restrict__ fS,
                                             const unsigned *__restrict__ row, const unsigned st_idx,
                                             const unsigned stp, const unsigned swp, const T tol, const T gst,
                                             const int vtrack, const int stats, PlainState<T> *__restrict__ st,
                                             PlainSlot *__restrict__ slots)
{
  // n is the length of the V columns (the table row has gridDim.x pairs)
  __shared__ double lds[8 * (B / 64) + 2];
  __shared__ int last;
  if (*(volatile int *)&st->done)
    return;
  const int k = blockIdx.x;
  const int p = (int)row[2 * k], q = (int)row[2 * k + 1];
  const int lane = threadIdx.x;
  Step<T> S;
  S.tol = tol;
  S.gst = gst;
  S.sG = st->sG_next;
  S.sV = st->sV_next;
  S.vtrack = vtrack;

  T *const Gp = Gr + (long long)p * ldG, *const Gq = Gr + (long long)q * ldG;
  T *const Gpi = CPLX ? Gi + (long long)p * ldG : nullptr, *const Gqi = CPLX ? Gi + (long long)q * ldG : nullptr;
  T *const Vp = Vr + (long long)p * ldV, *const Vq = Vr + (long long)q * ldV;
  T *const Vpi = CPLX ? Vi + (long long)p * ldV : nullptr, *const Vqi = CPLX ? Vi + (long long)q * ldV : nullptr;
  // the V pair in registers along with G (loaded at the start, so that its
  // latency overlaps the reductions) when it fits E chunks of B rows (E <= 8,
  // complex E <= 4)
  const bool vreg = !STREAM && (E * (CPLX ? 2 : 1) <= 8) && (n <= E * B);
  if (S.sV && !vreg) {
    for (int i = lane; i < n; i += B) {
      Vp[i] = ldexp_(Vp[i], S.sV);
      Vq[i] = ldexp_(Vq[i], S.sV);
      if constexpr (CPLX) {
        Vpi[i] = ldexp_(Vpi[i], S.sV);
        Vqi[i] = ldexp_(Vqi[i], S.sV);
      }
    }
  }
  T eSp, fSp, eSq, fSq;
  PairOut<T, CPLX> o;
  PairDec<T> D;
  if (vreg) {
    T xr[E], yr[E], xi[E], yi[E], vx[E], vy[E], vxi[E], vyi[E];
    const int nEV = (n + B - 1) / B;
    load_rows<T, CPLX, B, E>(Gp, Gq, Gpi, Gqi, m, 0, S.sG, lane, xr, yr, xi, yi);
    load_rows<T, CPLX, B, E>(Vp, Vq, Vpi, Vqi, n, 0, S.sV, lane, vx, vy, vxi, vyi);
    Step<T> Sg = S;
    Sg.vtrack = 0;
    o = pair_step<T, CPLX, B, E>(m, 0, xr, yr, xi, yi, nullptr, nullptr, nullptr, nullptr, eSp, fSp, eSq, fSq, Sg, lds,
                                 lane, D);
    if (!o.err && ((o.code != 1) || S.sG))
      store_rows<T, CPLX, B, E>(Gp, Gq, Gpi, Gqi, m, 0, lane, xr, yr, xi, yi);
    if (!D.err) {
      T tV = v_apply<T, CPLX, E>(D, vx, vy, vxi, vyi, nEV);
      if (vtrack && ((D.code == 2) || (D.code == -2)) && !((D.at == (T)0) && (!CPLX || (D.sat == (T)0)))) {
        tV = red_max<T, B>(tV, lds);
        o.tV = tV;
        if (!o.err && !isfinite_(tV))
          o.err = CPLX ? -638 : -588;
      }
    }
    if (S.sV || (!D.err && (D.code != 1) && (D.code != 3)))
      store_rows<T, CPLX, B, E>(Vp, Vq, Vpi, Vqi, n, 0, lane, vx, vy, vxi, vyi);
  }
  else if (!STREAM) {
    T xr[E], yr[E], xi[E], yi[E];
    load_rows<T, CPLX, B, E>(Gp, Gq, Gpi, Gqi, m, 0, S.sG, lane, xr, yr, xi, yi);
    o = pair_step<T, CPLX, B, E>(m, n, xr, yr, xi, yi, Vp, Vq, Vpi, Vqi, eSp, fSp, eSq, fSq, S, lds, lane, D);
    if (!o.err && ((o.code != 1) || S.sG))
      store_rows<T, CPLX, B, E>(Gp, Gq, Gpi, Gqi, m, 0, lane, xr, yr, xi, yi);
  }
  else {
    typedef typename Acc<T>::A A;
    const int nc = (m + B * E - 1) / (B * E); // chunks of E rows per lane
    T xr[E], yr[E], xi[E], yi[E];
    // The sums are taken in the same pass as the maxima, scaled by a
    // predicted exponent: the norm exponent each column was left with (every
    // step evaluates every column), moved by a pending rescaling.  Power-of-
    // two scaling commutes with every rounding of the sums as long as no
    // product or error term leaves the normal range under either exponent,
    // which holds when every nonzero element scaled by either exponent lies
    // in [2^-440, 2^100]; the sums are then rescaled to the max-exponent
    // scaling exactly, and otherwise taken again in a second pass.
    const T epx = eS[p], epy = eS[q];
    const bool pred = isfinite_(epx) && isfinite_(epy);
    const int kpx = pred ? (int)epx + S.sG : 0, kpy = pred ? (int)epy + S.sG : 0;
    T mx = (T)0, my = (T)0, nx = inf_<T>(), ny = inf_<T>(); // max and min nonzero |element|
    DD<A> sx{0, 0}, sy{0, 0}, dr{0, 0}, di{0, 0};
    for (int c = 0; c < nc; ++c) {
      load_rows<T, CPLX, B, E>(Gp, Gq, Gpi, Gqi, m, c * E, S.sG, lane, xr, yr, xi, yi);
      T cx, cy;
      pair_maxima<T, CPLX, E>(xr, yr, xi, yi, E, cx, cy);
      mx = __builtin_fmax(mx, cx);
      my = __builtin_fmax(my, cy);
      if (pred) {
        pair_sums<T, CPLX, E>(xr, yr, xi, yi, E, kpx, kpy, sx, sy, dr, di);
#pragma unroll
        for (int j = 0; j < E; ++j) {
          nx = __builtin_fmin(nx, (xr[j] != (T)0) ? vabs(xr[j]) : inf_<T>());
          ny = __builtin_fmin(ny, (yr[j] != (T)0) ? vabs(yr[j]) : inf_<T>());
          if constexpr (CPLX) {
            nx = __builtin_fmin(nx, (xi[j] != (T)0) ? vabs(xi[j]) : inf_<T>());
            ny = __builtin_fmin(ny, (yi[j] != (T)0) ? vabs(yi[j]) : inf_<T>());
          }
        }
      }
    }
    red_max2<T, B>(mx, my, lds);
    const int kx = max_exp(mx), ky = max_exp(my);
    bool one = false;
    if (pred) {
      T mnx = -nx, mny = -ny;
      red_max2<T, B>(mnx, mny, lds);
      auto fits = [](const T mn, const T mxv, const int k, const int kp) {
        if (!(mxv > (T)0))
          return true; // a zero column: every sum is zero under any scaling
        const int en = frexp_exp(-mn) - 1;
        const int hi = (k > kp) ? k : kp;
        return (en - hi >= -440) && (k - kp <= 100);
      };
      one = fits(mnx, mx, kx, kpx) && fits(mny, my, ky, kpy);
    }
    if (one) {
      red_dd4<A, B>(sx, sy, dr, di, lds);
      const int dx = kpx - kx, dy = kpy - ky;
      sx = DD<A>{ldexp_(sx.s, 2 * dx), ldexp_(sx.c, 2 * dx)};
      sy = DD<A>{ldexp_(sy.s, 2 * dy), ldexp_(sy.c, 2 * dy)};
      dr = DD<A>{ldexp_(dr.s, dx + dy), ldexp_(dr.c, dx + dy)};
      di = DD<A>{ldexp_(di.s, dx + dy), ldexp_(di.c, dx + dy)};
    }
    else {
      sx = sy = dr = di = DD<A>{0, 0};
      for (int c = 0; c < nc; ++c) {
        load_rows<T, CPLX, B, E>(Gp, Gq, Gpi, Gqi, m, c * E, S.sG, lane, xr, yr, xi, yi);
        pair_sums<T, CPLX, E>(xr, yr, xi, yi, E, kx, ky, sx, sy, dr, di);
      }
      red_dd4<A, B>(sx, sy, dr, di, lds);
    }
    const PairScal<T> P = pair_scal<T, CPLX>(mx, my, kx, ky, sx, sy, dr, di);
    D = pair_decide<T, CPLX>(P, S.tol, S.gst);
    T tG = (T)0;
    for (int c = 0; c < nc; ++c) {
      load_rows<T, CPLX, B, E>(Gp, Gq, Gpi, Gqi, m, c * E, S.sG, lane, xr, yr, xi, yi);
      // the V pair is updated with the first chunk
      o = pair_apply<T, CPLX, B, E, false>(c ? 0 : n, xr, yr, xi, yi, Vp, Vq, Vpi, Vqi, P, D, S.vtrack, eSp, fSp, eSq,
                                           fSq, lds, lane, E);
      if (o.err)
        break;
      tG = xmax(tG, o.tG);
      if ((o.code != 1) || S.sG)
        store_rows<T, CPLX, B, E>(Gp, Gq, Gpi, Gqi, m, c * E, lane, xr, yr, xi, yi);
    }
    if (!o.err) {
      tG = red_max<T, B>(tG, lds);
      o.tG = tG;
      if (((o.code == 2) || (o.code == -2) || (o.code == 3) || (o.code == -3)) && !isfinite_(tG))
        o.err = CPLX ? -630 : -580;
    }
  }
  if (lane == 0) {
    // this pair's report, then the ticket; the wait makes the slot words
    // complete before the ticket is taken
    PlainSlot *const sl = slots + blockIdx.x;
    __hip_atomic_store(&sl->err, o.err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (!o.err) {
      eS[p] = eSp; fS[p] = fSp;
      eS[q] = eSq; fS[q] = fSq;
      __hip_atomic_store(&sl->tG, (double)o.tG, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      __hip_atomic_store(&sl->tV, (double)o.tV, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      __hip_atomic_store(&sl->trans, (unsigned)o.trans, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      if (stats) {
        // |sin| of the rotation: c |tan|, the complex tangent being at + i sat
        T ta = vabs(D.at);
        if constexpr (CPLX)
          ta = sqrt_(fma_(D.sat, D.sat, ta * ta));
        const T sn = D.trans ? D.c * ta : (T)0;
        __hip_atomic_store(&sl->cosv, (double)D.cosv, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __hip_atomic_store(&sl->sinv, (double)sn, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      }
    }
    __builtin_amdgcn_s_waitcnt(0);
    last = (__hip_atomic_fetch_add(&st->blocks_done, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == gridDim.x - 1u);
  }
  __syncthreads();
  if (!last)
    return;
  // ---- last workgroup of the step: reduce the slots (all lanes), fold the
  // maxima, decide the next step's rescaling (dvjsvd.c:198-233) and the end
  // of the sweep (:599-601)
  double rG = 0.0, rV = 0.0, rc = 0.0, rs = 0.0;
  unsigned long long rt = 0ull, rkey = ~0ull; // first failing pair: (index, code)
  for (int i = lane; i < (int)gridDim.x; i += B) {
    const PlainSlot *const sl = slots + i;
    const int e = __hip_atomic_load(&sl->err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (e) {
      const unsigned long long key = ((unsigned long long)(unsigned)i << 32) | (unsigned)e;
      rkey = (key < rkey) ? key : rkey;
      continue;
    }
    rG = fmax(rG, __hip_atomic_load(&sl->tG, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
    rV = fmax(rV, __hip_atomic_load(&sl->tV, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
    rt += __hip_atomic_load(&sl->trans, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (stats) {
      rc = fmax(rc, __hip_atomic_load(&sl->cosv, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
      rs = fmax(rs, __hip_atomic_load(&sl->sinv, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
    }
  }
#pragma unroll
  for (int o2 = 32; o2 >= 1; o2 >>= 1) {
    rG = fmax(rG, __shfl_xor(rG, o2, 64));
    rV = fmax(rV, __shfl_xor(rV, o2, 64));
    rc = fmax(rc, __shfl_xor(rc, o2, 64));
    rs = fmax(rs, __shfl_xor(rs, o2, 64));
    rt += __shfl_xor(rt, o2, 64);
    const unsigned long long k2 = __shfl_xor(rkey, o2, 64);
    rkey = (k2 < rkey) ? k2 : rkey;
  }
  if constexpr (B > 64) {
    double *const w = lds + 6 * (lane >> 6);
    if ((lane & 63) == 0) {
      w[0] = rG; w[1] = rV; w[2] = rc; w[3] = rs;
      w[4] = __longlong_as_double((long long)rt);
      w[5] = __longlong_as_double((long long)rkey);
    }
    __syncthreads();
    for (int w2 = 1; w2 < B / 64; ++w2) {
      const double *const u = lds + 6 * w2;
      rG = fmax(rG, u[0]); rV = fmax(rV, u[1]); rc = fmax(rc, u[2]); rs = fmax(rs, u[3]);
      rt += (unsigned long long)__double_as_longlong(u[4]);
      const unsigned long long k2 = (unsigned long long)__double_as_longlong(u[5]);
      rkey = (k2 < rkey) ? k2 : rkey;
    }
  }
  if (lane)
    return;
  st->blocks_done = 0u;
  if (rkey != ~0ull) {
    st->status = (int)(unsigned)(rkey & 0xffffffffull);
    st->err_sw = st->sweeps;
    st->err_st = st_idx;
    st->done = 1;
    return;
  }
  const T nMG = (T)rG, nMV = (T)rV;
  st->swt += rt;
  if (stats) {
    st->cos_max = fmax(st->cos_max, rc);
    st->sin_max = fmax(st->sin_max, rs);
  }
  // st->MG already carries this step's shift (applied when it was decided)
  T MG = st->MG;
  MG = (nMG > MG) ? nMG : MG;
  T MV = st->MV;
  MV = (nMV > MV) ? nMV : MV;
  const int rot = CPLX ? Lim<T>::ROT_EXP_C - 1 : Lim<T>::ROT_EXP_R;
  int sG = 0, sV = 0;
  {
    const int eM = frexp_exp(MG) - 1;
    if (rot - eM < 0) {
      sG = st->NRM_EXP - eM - 1;
      MG = ldexp_(MG, sG);
      st->sT += sG;
      if (st->nev < 32u) {  // logged against the step it precedes
        const bool wrap = (st_idx + 1u == stp);
        auto &e = st->ev[st->nev++];
        e.sw = st->sweeps + (wrap ? 1u : 0u);
        e.st = wrap ? 0u : st_idx + 1u;
        e.eM = eM;
        e.sR = rot - eM;
        e.sN = sG;
        e.M = MG;
      }
    }
  }
  {
    const int eV = frexp_exp(MV) - 1;
    if (rot - eV < 0) {
      sV = rot - eV;
      MV = ldexp_(MV, sV);
      st->sV += sV;
    }
  }
  st->MG = MG;
  st->MV = MV;
  st->sG_next = sG;
  st->sV_next = sV;
  if (st_idx == stp - 1u) {
    if (stats) {
      st->last_cos = st->cos_max;
      st->last_sin = st->sin_max;
      st->cos_max = st->sin_max = 0.0;
    }
    const unsigned long long swt = st->swt;
    if (st->sweeps < 64u)
      st->hist[st->sweeps] = swt;
    if (!swt)
      st->done = 1;
    else {
      st->sweeps += 1u;
      st->swt = 0ull;
      if (st->sweeps >= swp)
        st->done = 1;
    }
  }
  __threadfence();
}

// ----------------------------------------------------------------- epilogue
template <typename T>
__global__ __launch_bounds__(256) void k_epilogue(const int m, const int n, T *__restrict__ Gr, T *__restrict__ Gi,
                                                  const long long ldG, T *__restrict__ Vr, T *__restrict__ Vi,
                                                  const long long ldV, T *__restrict__ eS, T *__restrict__ fS,
                                                  const PlainState<T> *__restrict__ st, const unsigned swp)
{
  // dvjsvd.c:604-631: G_j <- G_j / fS_j * 2^-eS_j, eS_j -= sT; V *= 2^-sV
  if (st->status || st->zero || !(st->sweeps < swp))
    return;
  const int j = blockIdx.x;
  const T f_ = fS[j], s_ = -eS[j];
  for (int i = threadIdx.x; i < m; i += blockDim.x) {
    Gr[j * ldG + i] = scalef(Gr[j * ldG + i] / f_, s_);
    if (Gi)
      Gi[j * ldG + i] = scalef(Gi[j * ldG + i] / f_, s_);
  }
  if (st->sV) {
    const T sv = (T)(-st->sV);
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
      Vr[j * ldV + i] = scalef(Vr[j * ldV + i], sv);
      if (Vi)
        Vi[j * ldV + i] = scalef(Vi[j * ldV + i], sv);
    }
  }
  __syncthreads();
  if (threadIdx.x == 0)
    eS[j] -= (T)st->sT;
}

// ------------------------------------------------------------------ driver
template <typename T, bool CPLX, int B, int E, bool STREAM = false>
static int launch_steps(const PlainArgs<T> &a, PlainState<T> *st, PlainSlot *slots, const unsigned *row, unsigned s,
                        const T tol, const T gst, hipStream_t strm)
{
  hipLaunchKernelGGL((k_step<T, CPLX, B, E, STREAM>), dim3(a.n / 2), dim3(B), 0, strm, (int)a.m,
                     (int)(a.nv ? a.nv : a.n), a.Gr, a.Gi, a.ldG, a.Vr, a.Vi, a.ldV, a.eS, a.fS, row, s, a.stp, a.swp,
                     tol, gst, a.vpreset, a.stats, st, slots);
  return 0;
}

template <typename T, bool CPLX>
using StepFn = int (*)(const PlainArgs<T> &, PlainState<T> *, PlainSlot *, const unsigned *, unsigned, T, T,
                       hipStream_t);

template <typename T, bool CPLX>
static StepFn<T, CPLX> pick_step(long long m, int &B_out)
{
  // smallest group that holds the column pair in registers with <= 8
  // elements per lane, then 512 lanes (two waves per SIMD, 256 registers per
  // lane) with 8 or 16; longer columns are streamed through 512 lanes in
  // chunks of 4 rows per lane (a 1024-lane group would cap the registers at
  // 128 and spill the pair)
  if (m <= 64) { B_out = 64; return &launch_steps<T, CPLX, 64, 1>; }
  if (m <= 128) { B_out = 64; return &launch_steps<T, CPLX, 64, 2>; }
  if (m <= 256) { B_out = 64; return &launch_steps<T, CPLX, 64, 4>; }
  if (m <= 512) { B_out = 64; return &launch_steps<T, CPLX, 64, 8>; }
  if (m <= 1024) { B_out = 256; return &launch_steps<T, CPLX, 256, 4>; }
  if (m <= 2048) { B_out = 256; return &launch_steps<T, CPLX, 256, 8>; }
  if (m <= 4096) { B_out = 512; return &launch_steps<T, CPLX, 512, 8>; }
  if constexpr (!CPLX) {
    if (m <= 8192) { B_out = 512; return &launch_steps<T, CPLX, 512, 16>; }
  }
  if (m <= 0x7fffffffLL / 2) { B_out = 512; return &launch_steps<T, CPLX, 512, 4, true>; }
  return nullptr;
}

template <typename T> static T ldexp_host(T x, int e) { return std::ldexp(x, e); }

template <typename T, bool CPLX>
int plain_svd(const PlainArgs<T> &a, PlainRun *run, hipStream_t s)
{
  if (a.n == 0)
    return 0;
  int B = 0;
  StepFn<T, CPLX> step = pick_step<T, CPLX>(a.m, B);
  if (!step)
    return set_error(-1, "m = %lld exceeds the largest supported column length", a.m);
  // control block and per-pair slots, released in stream order on return
  Staging sg(s);
  PlainState<T> *st = nullptr;
  PlainSlot *slots = nullptr;
  if (int rc = sg.alloc(&st, 1))
    return rc;
  if (int rc = sg.alloc(&slots, (size_t)(a.n / 2)))
    return rc;
  VJ_HIP(hipMemsetAsync(st, 0, sizeof(PlainState<T>), s));
  const int mx_blocks = 1024;
  hipLaunchKernelGGL(k_maxabs<T>, dim3(mx_blocks), dim3(256), 0, s, (int)a.m, (int)a.n, a.Gr, a.ldG, &st->MG_u);
  if (CPLX)
    hipLaunchKernelGGL(k_maxabs<T>, dim3(mx_blocks), dim3(256), 0, s, (int)a.m, (int)a.n, a.Gi, a.ldG, &st->MG_u);
  const long long nv = a.nv ? a.nv : a.n;
  if (a.vpreset) {
    hipLaunchKernelGGL(k_maxabs<T>, dim3(mx_blocks), dim3(256), 0, s, (int)nv, (int)a.n, a.Vr, a.ldV, &st->MV_u);
    if (CPLX)
      hipLaunchKernelGGL(k_maxabs<T>, dim3(mx_blocks), dim3(256), 0, s, (int)nv, (int)a.n, a.Vi, a.ldV, &st->MV_u);
  }
  hipLaunchKernelGGL((k_setup<T, CPLX>), dim3(1), dim3(64), 0, s, (int)a.m, st, a.vpreset);
  hipLaunchKernelGGL(k_init_cols<T>, dim3((unsigned)a.n), dim3(256), 0, s, (int)a.m, (int)nv, a.Gr, a.Gi, a.ldG,
                     a.Vr, a.Vi, a.ldV, a.eS, a.fS, st, a.vpreset);
  if (int rc = check_launch())
    return rc;
  const T tol = (a.tol > (T)0) ? a.tol : host_tol<T>(a.m);
  const T gst = host_gst<T>(tol);
  // Sweeps are enqueued one ahead of the convergence check: the host only
  // reads the device flag of sweep k-1 after sweep k has been queued.
  PlainFlags *hf = nullptr;
  VJ_HIP(hipHostMalloc((void **)&hf, 2 * sizeof(PlainFlags), hipHostMallocDefault));
  hipEvent_t ev[2];
  VJ_HIP(hipEventCreateWithFlags(&ev[0], hipEventDisableTiming));
  VJ_HIP(hipEventCreateWithFlags(&ev[1], hipEventDisableTiming));
  int rc = 0;
  unsigned launched = 0;
  hipEvent_t tev[2] = {nullptr, nullptr};
  if (run && (hipEventCreate(&tev[0]) != hipSuccess || hipEventCreate(&tev[1]) != hipSuccess ||
              hipEventRecord(tev[0], s) != hipSuccess))
    rc = set_error(VJ_EHIP, "event creation failed");
  for (unsigned sw = 0; !rc && sw < a.swp; ++sw) {
    for (unsigned k = 0; k < a.stp; ++k)
      step(a, st, slots, a.js + (size_t)k * a.n, k, tol, gst, s);
    if ((rc = check_launch()))
      break;
    ++launched;
    if (hipMemcpyAsync(&hf[sw & 1u], &st->done, sizeof(PlainFlags), hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipEventRecord(ev[sw & 1u], s) != hipSuccess) {
      rc = set_error(VJ_EHIP, "flag copy failed");
      break;
    }
    if (sw >= 1) {
      if (hipEventSynchronize(ev[(sw - 1) & 1u]) != hipSuccess) {
        rc = set_error(VJ_EHIP, "event sync failed");
        break;
      }
      if (hf[(sw - 1) & 1u].done)
        break;
    }
  }
  if (!rc && !a.raw) {
    hipLaunchKernelGGL(k_epilogue<T>, dim3((unsigned)a.n), dim3(256), 0, s, (int)a.m, (int)nv, a.Gr, a.Gi, a.ldG,
                       a.Vr, a.Vi, a.ldV, a.eS, a.fS, st, a.swp);
    rc = check_launch();
  }
  if (!rc && run && hipEventRecord(tev[1], s) != hipSuccess)
    rc = set_error(VJ_EHIP, "event record failed");
  PlainState<T> h;
  if (!rc) {
    if (hipMemcpyAsync(&h, st, sizeof(h), hipMemcpyDeviceToHost, s) != hipSuccess || hipStreamSynchronize(s) != hipSuccess)
      rc = set_error(VJ_EHIP, "state copy failed");
  }
  float steps_ms = 0.0f;
  if (!rc && run && hipEventElapsedTime(&steps_ms, tev[0], tev[1]) != hipSuccess)
    steps_ms = -1.0f;
  for (int i = 0; i < 2; ++i)
    if (tev[i])
      (void)hipEventDestroy(tev[i]);
  (void)hipEventDestroy(ev[0]);
  (void)hipEventDestroy(ev[1]);
  (void)hipHostFree(hf);
  if (rc)
    return rc;
  if (run) {
    run->sweeps = h.sweeps;
    run->sT = h.sT;
    run->sV = h.sV;
    run->launched_sweeps = launched;
    for (int i = 0; i < 64; ++i)
      run->hist[i] = h.hist[i];
    run->MG0 = (double)h.MG0;
    run->MG_scaled = (double)ldexp_host(h.MG0, h.sG_init);
    run->MG_final = (double)h.MG;
    run->eM0 = h.eM0;
    run->sR0 = h.sR0;
    run->sN0 = h.sG_init;
    run->zero = h.zero;
    run->status = h.status;
    run->nev = h.nev;
    for (unsigned i = 0; i < h.nev && i < 32u; ++i) {
      run->ev[i].sw = h.ev[i].sw;
      run->ev[i].st = h.ev[i].st;
      run->ev[i].eM = h.ev[i].eM;
      run->ev[i].sR = h.ev[i].sR;
      run->ev[i].sN = h.ev[i].sN;
      run->ev[i].M = (double)h.ev[i].M;
    }
    run->err_sw = h.err_sw;
    run->err_st = h.err_st;
    run->steps_ms = steps_ms;
    run->last_cos = h.last_cos;
    run->last_sin = h.last_sin;
    run->sG_pending = h.sG_next;
    run->sV_pending = h.sV_next;
  }
  if (h.status)
    return h.status;
  if (h.zero)
    return 0;
  return (int)h.sweeps;
}

template int plain_svd<double, false>(const PlainArgs<double> &, PlainRun *, hipStream_t);
template int plain_svd<double, true>(const PlainArgs<double> &, PlainRun *, hipStream_t);
template int plain_svd<float, false>(const PlainArgs<float> &, PlainRun *, hipStream_t);
template int plain_svd<float, true>(const PlainArgs<float> &, PlainRun *, hipStream_t);

} // namespace vj
