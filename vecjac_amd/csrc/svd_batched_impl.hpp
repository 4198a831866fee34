// Batched one-sided Jacobi SVD of many small matrices of uneven sizes in one
// persistent launch (no reference counterpart; each matrix follows
// dvjsvd_/zvjsvd_, dvjsvd.c:40-638, with V := I).
//
// Workgroups pull matrices from a queue sorted by decreasing cost (m n^2),
// so the long ones start first and the short ones fill the tail.  A
// workgroup of W waves owns one matrix at a time and runs every sweep of it
// without leaving the kernel.  The pairs of a step (the modulus ordering,
// computed, not read from a table) are dealt to the waves; a wave takes its
// pairs two at a time (wave_step): the next two are loaded while these are
// reduced (norms and dot products with DPP / permlane wave reductions), lane
// r decides pair r (convergence test, 2x2 EVD), and the decisions are
// applied to G and V in registers.  One workgroup barrier ends a step; the
// per-matrix control state (max-|.| bounds, shifts, transformation counts)
// is folded through double-buffered LDS slots, so every thread takes the
// same rescaling and convergence decisions; no host round trip.  The
// register footprint (64-row chunks per lane) is chosen per matrix.
//
// Short columns (m <= 128 real, 64 complex) go to a second launch of the
// same kernel in its lane-group form (group_step): four pairs per wave, one
// per row of 16 lanes, so that the reductions, the decisions and their
// scalar work are shared by four pairs instead of two.  The two forms are
// separate launches because one body holding both exceeds 256 registers.
//
// The templates are instantiated once per precision, in svd_batched_d.hip,
// svd_batched_z.hip, svd_batched_s.hip and svd_batched_c.hip, so that the
// four compile in parallel.
#pragma once

#include "svd_common.hpp"
#include "svd_batched.hpp"
#include "pjs.hpp"

namespace vj {

// Column pair (p, q) of an m-row matrix into registers, row lane + 64 j.
// All E chunks are loaded unconditionally (rows past m read row m - 1 and
// are zeroed at first use): a fixed number of loads per pair lets the
// compiler wait for exactly the loads of the pair in use and not for the
// prefetched ones behind them.
template <typename T, bool CPLX, int E>
__device__ __forceinline__ void load_pair(const T *Gr, const T *Gi, const int m, const int nE, const int p, const int q,
                                          const int lane, T (&xr)[E], T (&yr)[E], T (&xi)[E], T (&yi)[E])
{
  (void)nE;
  const long long op = (long long)p * m, oq = (long long)q * m;
#pragma unroll
  for (int j = 0; j < E; ++j) {
    const int i = lane + j * 64;
    const int ic = (i < m) ? i : (m - 1);
    xr[j] = Gr[op + ic];
    yr[j] = Gr[oq + ic];
    if constexpr (CPLX) {
      xi[j] = Gi[op + ic];
      yi[j] = Gi[oq + ic];
    }
  }
}

// Zero the rows past m and apply a pending 2^sG rescaling (at first use of
// a prefetched pair).
template <typename T, bool CPLX, int E>
__device__ __forceinline__ void ready_pair(const int m, const int nE, const int sG, const int lane, T (&xr)[E], T (&yr)[E],
                                           T (&xi)[E], T (&yi)[E])
{
#pragma unroll
  for (int j = 0; j < E; ++j) {
    if (j >= nE)
      break;
    if ((j == nE - 1) && (m & 63)) {  // only the last chunk can be partial
      const bool in = (lane + j * 64) < m;
      xr[j] = in ? xr[j] : (T)0;
      yr[j] = in ? yr[j] : (T)0;
      if constexpr (CPLX) {
        xi[j] = in ? xi[j] : (T)0;
        yi[j] = in ? yi[j] : (T)0;
      }
    }
    if constexpr (!CPLX)
      xi[j] = yi[j] = (T)0;
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

template <typename T, bool CPLX, int E>
__device__ __forceinline__ void store_pair(T *Gr, T *Gi, const int m, const int nE, const int p, const int q, const int lane,
                                           const T (&xr)[E], const T (&yr)[E], const T (&xi)[E], const T (&yi)[E])
{
  T *const Gp = Gr + (long long)p * m, *const Gq = Gr + (long long)q * m;
#pragma unroll
  for (int j = 0; j < E; ++j) {
    if (j >= nE)
      break;
    const int i = lane + j * 64;
    if (i < m) {
      Gp[i] = xr[j];
      Gq[i] = yr[j];
      if constexpr (CPLX) {
        Gi[(long long)p * m + i] = xi[j];
        Gi[(long long)q * m + i] = yi[j];
      }
    }
  }
}

// V column pair (p, q) of an n x n matrix (n <= 128) into registers, row
// lane + 64 j; unconditional loads as in load_pair.
template <typename T, bool CPLX>
__device__ __forceinline__ void load_vpair(const T *Vr, const T *Vi, const int n, const int p, const int q, const int lane,
                                           T (&vx)[2], T (&vy)[2], T (&vxi)[2], T (&vyi)[2])
{
  const long long op = (long long)p * n, oq = (long long)q * n;
#pragma unroll
  for (int j = 0; j < 2; ++j) {
    const int i = lane + 64 * j;
    const int ic = (i < n) ? i : (n - 1);
    vx[j] = Vr[op + ic];
    vy[j] = Vr[oq + ic];
    if constexpr (CPLX) {
      vxi[j] = Vi[op + ic];
      vyi[j] = Vi[oq + ic];
    }
    else
      vxi[j] = vyi[j] = (T)0;
  }
}

template <typename T>
__device__ __forceinline__ void scale_vpair(const int sV, T (&vx)[2], T (&vy)[2], T (&vxi)[2], T (&vyi)[2])
{
#pragma unroll
  for (int j = 0; j < 2; ++j) {
    vx[j] = ldexp_(vx[j], sV);
    vy[j] = ldexp_(vy[j], sV);
    vxi[j] = ldexp_(vxi[j], sV);
    vyi[j] = ldexp_(vyi[j], sV);
  }
}

template <typename T, bool CPLX>
__device__ __forceinline__ void store_vpair(T *Vr, T *Vi, const int n, const int p, const int q, const int lane,
                                            const T (&vx)[2], const T (&vy)[2], const T (&vxi)[2], const T (&vyi)[2])
{
#pragma unroll
  for (int j = 0; j < 2; ++j) {
    const int i = lane + 64 * j;
    if (i < n) {
      Vr[(long long)p * n + i] = vx[j];
      Vr[(long long)q * n + i] = vy[j];
      if constexpr (CPLX) {
        Vi[(long long)p * n + i] = vxi[j];
        Vi[(long long)q * n + i] = vyi[j];
      }
    }
  }
}

// The V part of pair_apply for V held in registers: the same swap or
// rotation, with the same fused multiply-adds.  With TRK, returns the lane's
// max |updated element| of a rotation (the tracking of a preset V, as
// v_apply); otherwise 0, and the maxima are not formed.
template <typename T, bool CPLX, bool TRK = false>
__device__ __forceinline__ T v_update(const PairDec<T> &D, T (&vx)[2], T (&vy)[2], T (&vxi)[2], T (&vyi)[2])
{
  T tV = (T)0;
  const int code = D.code;
  const T c = D.c, at = D.at, sat = D.sat;
  bool swap = (code == -1) || (code == -3);
  if ((code == 2) || (code == -2)) {
    const bool perm = (code < 0);
    if ((at == (T)0) && (!CPLX || (sat == (T)0)))
      swap = (c == (T)1) && perm;
    else {
      const bool c1 = (c == (T)1);
      const T cna = -at;
#pragma unroll
      for (int j = 0; j < 2; ++j) {
        const T x_r = vx[j], y_r = vy[j];
        if (!CPLX || (sat == (T)0)) {
          T a = perm ? fma_(-x_r, at, y_r) : fma_(y_r, at, x_r);
          T b = perm ? fma_(y_r, at, x_r) : fma_(-x_r, at, y_r);
          if (!c1) { a = a * c; b = b * c; }
          vx[j] = a;
          vy[j] = b;
          if constexpr (TRK)
            tV = max1p(tV, max1p(vabs(a), vabs(b)));
          if constexpr (CPLX) {
            const T x_i = vxi[j], y_i = vyi[j];
            T a2 = perm ? fma_(-x_i, at, y_i) : fma_(y_i, at, x_i);
            T b2 = perm ? fma_(y_i, at, x_i) : fma_(-x_i, at, y_i);
            if (!c1) { a2 = a2 * c; b2 = b2 * c; }
            vxi[j] = a2;
            vyi[j] = b2;
            if constexpr (TRK)
              tV = max1p(tV, max1p(vabs(a2), vabs(b2)));
          }
        }
        else if constexpr (CPLX) {
          const T x_i = vxi[j], y_i = vyi[j];
          T ar = fma_(y_r, at, fma_(-y_i, sat, x_r));
          T ai = fma_(y_r, sat, fma_(y_i, at, x_i));
          T br = fma_(x_r, cna, fma_(-x_i, sat, y_r));
          T bi = fma_(x_r, sat, fma_(x_i, cna, y_i));
          if (!c1) { ar = ar * c; ai = ai * c; br = br * c; bi = bi * c; }
          if constexpr (TRK)
            tV = xmax(tV, xmax(xmax(vabs(ar), vabs(ai)), xmax(vabs(br), vabs(bi))));
          if (perm) { vx[j] = br; vxi[j] = bi; vy[j] = ar; vyi[j] = ai; }
          else { vx[j] = ar; vxi[j] = ai; vy[j] = br; vyi[j] = bi; }
        }
      }
    }
  }
  if (swap) {
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      T t = vx[j]; vx[j] = vy[j]; vy[j] = t;
      if constexpr (CPLX) { t = vxi[j]; vxi[j] = vyi[j]; vyi[j] = t; }
    }
  }
  return tV;
}

// Pairs a wave works on together (their reductions share the cross-lane
// exchanges and their 2x2 decisions run side by side on different lanes)
// and whether the next group is prefetched into a second register set,
// bounded by the registers per column (rc): two prefetched pairs up to 8,
// two pairs without prefetch up to 16 (and E <= 8), else one prefetched
// pair.
template <typename T, bool CPLX, int E> struct StepShape {
  static constexpr int rc = E * (CPLX ? 2 : 1) * (int)(sizeof(T) / 4);
  static constexpr int GP = ((rc <= 16) && (E <= 8)) ? 2 : 1;
  static constexpr bool PF = (rc <= 8) || (GP == 1);
};

// Max of four values over the wave, every lane receiving all four totals.
template <typename T> __device__ __forceinline__ void red_max4(T &a, T &b, T &c, T &d)
{
  T a1 = a, c1 = c, b1 = b, d1 = d;
  swap_<true>(a1, c1);   // lanes < 32: a, lanes >= 32: c
  swap_<true>(b1, d1);
  T k0 = xmax(a1, c1), k1 = xmax(b1, d1);
  swap_<false>(k0, k1);  // rows (a, b, c, d)
  const T k = row_max(xmax(k0, k1));
  bcast_rows4(k, a, b, c, d);
}

// Pair r's scalars and decision from lane r, as wave-uniform values (so
// that the update loops branch once on them instead of per element).
template <typename T> __device__ __forceinline__ PairScal<T> lane_scal(const PairScal<T> &P, const int r)
{
  return PairScal<T>{readlane_(P.ex, r), readlane_(P.fx, r), readlane_(P.ey, r), readlane_(P.fy, r),
                     readlane_(P.a21r, r), readlane_(P.a21i, r)};
}
template <typename T> __device__ __forceinline__ PairDec<T> lane_dec(const PairDec<T> &D, const int r)
{
  // code (-3..3), trans (0/1) and err (0 or minus a line number) in one word
  const int pk = (D.code + 8) | (D.trans << 4) | ((-D.err) << 5);
  const int u = __builtin_amdgcn_readlane(pk, r);
  return PairDec<T>{(u & 15) - 8, (u >> 4) & 1, -(u >> 5), readlane_(D.c, r), readlane_(D.at, r), readlane_(D.sat, r)};
}

// One step of the strategy for the pairs of one wave (pairs wave, wave + W,
// wave + 2W, ... of step st of ordering ord, pjs id 4 or 6), GP pairs at a time:
// load (the next group is prefetched), reduce the GP pairs together, decide
// with lane r of every quad working on pair r, broadcast each decision and
// apply it in registers, store.  Accumulates this wave's per-lane max of the
// updated G elements, its transformation count and its first error.
template <typename T, bool CPLX, int E, int W, int GP, bool PF>
__device__ __forceinline__ void wave_step(T *Gr, T *Gi, T *Vr, T *Vi, T *eS, T *fS, const int m, const int n, const int nE,
                                          const int ord, const unsigned st, const int wave, const int lane,
                                          const Step<T> &S, T &wG, T &wV, unsigned long long &wcnt, int &werr,
                                          int *mod, const int gstep, unsigned long long &wev, unsigned long long &wxf)
{
  typedef typename Acc<T>::A A;
  constexpr int B = 64;
  constexpr unsigned long long PL = CPLX ? 2ull : 1ull; // planes
  const int npairs = n >> 1;
  if (wave >= npairs)
    return;
  const int cnt0 = (npairs - wave + W - 1) / W;      // pair slots of this wave (slot i: pair wave + W i)
  // A pair whose two columns were last modified before its previous
  // evaluation, one sweep (n steps) ago, would repeat that evaluation's
  // no-op (same columns, same decision: no transformation, the same norms),
  // so it is not evaluated again unless a rescaling is pending.  Lane j of
  // `slot` holds the j-th slot to evaluate.
  int cnt = cnt0, slot = lane;
  if ((S.sG == 0) && (S.sV == 0) && (gstep > n)) {
    bool need = false;
    if (lane < cnt0) {
      unsigned pp, qq;
      step_pair(ord, (unsigned)n, st, (unsigned)(wave + W * lane), pp, qq);
      need = (mod[pp] >= gstep - n) || (mod[qq] >= gstep - n);
    }
    const unsigned long long mk = __ballot(need);
    cnt = __popcll(mk);
    if (cnt == 0)
      return;
    const int below = __popcll(mk & ((1ull << lane) - 1ull));
    const int pos = need ? below : (cnt + lane - below); // a permutation of the lanes
    slot = __builtin_amdgcn_ds_permute(pos << 2, lane);
  }
  wev += (unsigned long long)cnt;
  wxf += (unsigned long long)cnt * 2ull * PL * (unsigned long long)(m + n);
  const int ngr = (cnt + GP - 1) / GP;
  T xr[GP][E], yr[GP][E], xi[GP][E], yi[GP][E], vx[GP][2], vy[GP][2], vxi[GP][2], vyi[GP][2];
  unsigned p[GP], q[GP];
  // pair r of group g (a group past the end repeats the last pair)
  auto pair_of = [&](const int g, const int r, unsigned &pp, unsigned &qq) {
    int i = g * GP + r;
    i = (i < cnt) ? i : (cnt - 1);
    i = __builtin_amdgcn_readlane(slot, i);
    step_pair(ord, (unsigned)n, st, (unsigned)(wave + W * i), pp, qq);
  };
  if constexpr (PF) {
#pragma unroll
    for (int r = 0; r < GP; ++r) {
      pair_of(0, r, p[r], q[r]);
      load_vpair<T, CPLX>(Vr, Vi, n, (int)p[r], (int)q[r], lane, vx[r], vy[r], vxi[r], vyi[r]);
      load_pair<T, CPLX, E>(Gr, Gi, m, nE, (int)p[r], (int)q[r], lane, xr[r], yr[r], xi[r], yi[r]);
    }
  }
  for (int g = 0; g < ngr; ++g) {
    // prefetch of the next group (this group again after the last one), or
    // the loads of this group
    const int gn = PF ? ((g + 1 < ngr) ? (g + 1) : g) : g;
    unsigned pn[GP], qn[GP];
    T nxr[GP][E], nyr[GP][E], nxi[GP][E], nyi[GP][E], nvx[GP][2], nvy[GP][2], nvxi[GP][2], nvyi[GP][2];
#pragma unroll
    for (int r = 0; r < GP; ++r) {
      if constexpr (PF) {
        pair_of(gn, r, pn[r], qn[r]);
        load_vpair<T, CPLX>(Vr, Vi, n, (int)pn[r], (int)qn[r], lane, nvx[r], nvy[r], nvxi[r], nvyi[r]);
        load_pair<T, CPLX, E>(Gr, Gi, m, nE, (int)pn[r], (int)qn[r], lane, nxr[r], nyr[r], nxi[r], nyi[r]);
      }
      else {
        pair_of(g, r, p[r], q[r]);
        load_vpair<T, CPLX>(Vr, Vi, n, (int)p[r], (int)q[r], lane, vx[r], vy[r], vxi[r], vyi[r]);
        load_pair<T, CPLX, E>(Gr, Gi, m, nE, (int)p[r], (int)q[r], lane, xr[r], yr[r], xi[r], yi[r]);
      }
    }
    // ---- reduce the GP pairs together
    T mx[GP], my[GP];
    int kx[GP], ky[GP];
    DD<A> sx[GP], sy[GP], dr[GP], di[GP];
#pragma unroll
    for (int r = 0; r < GP; ++r) {
      ready_pair<T, CPLX, E>(m, nE, S.sG, lane, xr[r], yr[r], xi[r], yi[r]);
      pair_maxima<T, CPLX, E>(xr[r], yr[r], xi[r], yi[r], nE, mx[r], my[r]);
    }
    if constexpr (GP == 1)
      red_max2<T, B>(mx[0], my[0], nullptr);
    else {
#pragma unroll
      for (int r = 0; r < GP; r += 2)
        red_max4(mx[r], my[r], mx[r + 1], my[r + 1]);
    }
#pragma unroll
    for (int r = 0; r < GP; ++r) {
      kx[r] = max_exp(mx[r]);
      ky[r] = max_exp(my[r]);
      sx[r] = sy[r] = dr[r] = di[r] = DD<A>{0, 0};
      pair_sums<T, CPLX, E>(xr[r], yr[r], xi[r], yi[r], nE, kx[r], ky[r], sx[r], sy[r], dr[r], di[r]);
      red_dd4<A, B>(sx[r], sy[r], dr[r], di[r], nullptr);
    }
    // ---- lane r (mod GP) finishes and decides pair r
    const int rl = lane & (GP - 1);
    T lmx = mx[0], lmy = my[0];
    int lkx = kx[0], lky = ky[0];
    DD<A> lsx = sx[0], lsy = sy[0], ldr = dr[0], ldi = di[0];
#pragma unroll
    for (int r = 1; r < GP; ++r) {
      const bool t = (rl == r);
      lmx = t ? mx[r] : lmx;
      lmy = t ? my[r] : lmy;
      lkx = t ? kx[r] : lkx;
      lky = t ? ky[r] : lky;
      lsx = dd_sel(t, sx[r], lsx);
      lsy = dd_sel(t, sy[r], lsy);
      ldr = dd_sel(t, dr[r], ldr);
      ldi = dd_sel(t, di[r], ldi);
    }
    const PairScal<T> Pl = pair_scal<T, CPLX>(lmx, lmy, lkx, lky, lsx, lsy, ldr, ldi);
    const PairDec<T> Dl = pair_decide<T, CPLX>(Pl, S.tol, S.gst);
    // ---- apply and store, pair by pair
#pragma unroll
    for (int r = 0; r < GP; ++r) {
      if (g * GP + r >= cnt)
        break;
      const PairDec<T> D = lane_dec(Dl, r);
      // the scalars are needed on every lane only by a Gram-Schmidt update;
      // otherwise they only give the (e,f) outputs, written by lane r
      // (which holds pair r's own values)
      const bool gsr = (D.code == 3) || (D.code == -3);
      const PairScal<T> P = gsr ? lane_scal(Pl, r) : Pl;
      T eSp, fSp, eSq, fSq;
      const PairOut<T, CPLX> o = pair_apply<T, CPLX, B, E, false>(0, xr[r], yr[r], xi[r], yi[r], nullptr, nullptr, nullptr,
                                                                  nullptr, P, D, 0, eSp, fSp, eSq, fSq, nullptr, lane, nE);
      // an update that overflowed anywhere in the wave (the reference's
      // max-|.| test after the update) leaves G unstored
      int err = D.err;
      if (!err && (__ballot(!isfinite_(o.tG)) != 0ull))
        err = CPLX ? -630 : -580;
      if (!err && ((D.code != 1) || S.sG)) {
        store_pair<T, CPLX, E>(Gr, Gi, m, nE, (int)p[r], (int)q[r], lane, xr[r], yr[r], xi[r], yi[r]);
        if (lane == 0)
          mod[p[r]] = mod[q[r]] = gstep;
        wxf += 2ull * PL * (unsigned long long)m;
      }
      T tV = (T)0;
      if (S.sV || (!D.err && (D.code != 1) && (D.code != 3))) {
        if (S.sV)
          scale_vpair<T>(S.sV, vx[r], vy[r], vxi[r], vyi[r]);
        if (!D.err) {
          if (S.vtrack)
            tV = v_update<T, CPLX, true>(D, vx[r], vy[r], vxi[r], vyi[r]);
          else
            v_update<T, CPLX>(D, vx[r], vy[r], vxi[r], vyi[r]);
        }
        store_vpair<T, CPLX>(Vr, Vi, n, (int)p[r], (int)q[r], lane, vx[r], vy[r], vxi[r], vyi[r]);
        wxf += 2ull * PL * (unsigned long long)n;
      }
      // a preset V is tracked as djrot_ tracks it (an overflowed update fails)
      if (S.vtrack && !err && (__ballot(!isfinite_(tV)) != 0ull))
        err = CPLX ? -638 : -588;
      if (err)
        werr = err;
      else {
        if (lane == r) {
          eS[p[r]] = eSp; fS[p[r]] = fSp;
          eS[q[r]] = eSq; fS[q[r]] = fSq;
        }
        wG = xmax(wG, o.tG);
        if (S.vtrack)
          wV = xmax(wV, tV);
        wcnt += (unsigned long long)o.trans;
      }
    }
    if constexpr (!PF)
      continue;
#pragma unroll
    for (int r = 0; r < GP; ++r) {
      p[r] = pn[r];
      q[r] = qn[r];
#pragma unroll
      for (int j = 0; j < E; ++j) {
        xr[r][j] = nxr[r][j]; yr[r][j] = nyr[r][j]; xi[r][j] = nxi[r][j]; yi[r][j] = nyi[r][j];
      }
#pragma unroll
      for (int j = 0; j < 2; ++j) {
        vx[r][j] = nvx[r][j]; vy[r][j] = nvy[r][j]; vxi[r][j] = nvxi[r][j]; vyi[r][j] = nvyi[r][j];
      }
    }
  }
}

// ------------------------------------------------------------ block rounds
// Block form of the wave steps (real data, block modulus ordering, pjs id
// 6): a wave takes whole units of a round (a block pair: four columns; or a
// block met by itself: two), loads their G and V columns once, runs both
// steps of the round on them in registers (the two pairs of a step reduced
// and decided together, as in wave_step), and stores each changed column
// once: half the column traffic and half the barriers of two wave_step
// calls, with the results of those two steps.  (Between the two steps of a
// round the workgroup folds nothing: the fold only acts through a rescaling
// of G or V, which cannot occur in a batched run of real data, see k_bsvd.)

// Column c of a rows-row matrix from registers (row lane + 64 j).
template <typename T, int E>
__device__ __forceinline__ void store_col(T *A, const int rows, const int nE, const unsigned c, const int lane,
                                          const T (&x)[E])
{
  T *const Ac = A + (long long)c * rows;
#pragma unroll
  for (int j = 0; j < E; ++j) {
    if (j >= nE)
      break;
    const int i = lane + j * 64;
    if (i < rows)
      Ac[i] = x[j];
  }
}

// The two pairs (x0, y0) = columns (p0, q0) and (x1, y1) = (p1, q1) of one
// step, reduced and decided together (lane r & 1 decides pair r) and updated
// in place with their V columns; pair 1 is only applied when `two`.  Returns
// bit r (r = 0, 1) set when pair r changed its G columns, bit 2 + r when it
// changed its V columns; an error leaves everything unchanged from that pair
// on and is returned in werr.
template <typename T, int E>
__device__ __forceinline__ unsigned pair2(T (&x0)[E], T (&y0)[E], T (&x1)[E], T (&y1)[E], T (&v0)[2], T (&w0)[2],
                                          T (&v1)[2], T (&w1)[2], const bool two, const unsigned p0, const unsigned q0,
                                          const unsigned p1, const unsigned q1, const int nE, const int lane,
                                          const Step<T> &S, T *eS, T *fS, T &wG, unsigned long long &wcnt, int &werr)
{
  typedef typename Acc<T>::A A;
  constexpr int B = 64;
  T zi[E], vzi[2]; // imaginary parts of real data
#pragma unroll
  for (int j = 0; j < E; ++j)
    zi[j] = (T)0;
  vzi[0] = vzi[1] = (T)0;
  T mx0, my0, mx1, my1;
  pair_maxima<T, false, E>(x0, y0, zi, zi, nE, mx0, my0);
  pair_maxima<T, false, E>(x1, y1, zi, zi, nE, mx1, my1);
  red_max4(mx0, my0, mx1, my1);
  const int kx0 = max_exp(mx0), ky0 = max_exp(my0), kx1 = max_exp(mx1), ky1 = max_exp(my1);
  // six sums (the imaginary parts of the dots are zero): two reduced while
  // the second pair's are formed, then four
  DD<A> sx0{0, 0}, sy0{0, 0}, dr0{0, 0}, di0{0, 0}, sx1{0, 0}, sy1{0, 0}, dr1{0, 0}, di1{0, 0};
  pair_sums<T, false, E>(x0, y0, zi, zi, nE, kx0, ky0, sx0, sy0, dr0, di0);
  red_dd2<A>(sx0, sy0);
  pair_sums<T, false, E>(x1, y1, zi, zi, nE, kx1, ky1, sx1, sy1, dr1, di1);
  red_dd4<A, B>(dr0, sx1, sy1, dr1, nullptr);
  const bool t1 = (lane & 1) != 0;
  const PairScal<T> Pl = pair_scal<T, false>(t1 ? mx1 : mx0, t1 ? my1 : my0, t1 ? kx1 : kx0, t1 ? ky1 : ky0,
                                             dd_sel(t1, sx1, sx0), dd_sel(t1, sy1, sy0), dd_sel(t1, dr1, dr0),
                                             dd_sel(t1, di1, di0));
  const PairDec<T> Dl = pair_decide<T, false>(Pl, S.tol, S.gst);
  unsigned ch = 0u;
  auto apply = [&](T (&x)[E], T (&y)[E], T (&v)[2], T (&w)[2], const int r, const unsigned p, const unsigned q) {
    const PairDec<T> D = lane_dec(Dl, r);
    const bool gsr = (D.code == 3) || (D.code == -3);
    const PairScal<T> P = gsr ? lane_scal(Pl, r) : Pl;
    T eSp, fSp, eSq, fSq;
    const PairOut<T, false> o = pair_apply<T, false, B, E, false>(0, x, y, zi, zi, nullptr, nullptr, nullptr, nullptr,
                                                                  P, D, 0, eSp, fSp, eSq, fSq, nullptr, lane, nE);
    int err = D.err;
    if (!err && (__ballot(!isfinite_(o.tG)) != 0ull))
      err = -580;
    if (err) {
      werr = err;
      return;
    }
    if (D.code != 1)
      ch |= 1u << r;
    if ((D.code != 1) && (D.code != 3)) {
      v_update<T, false>(D, v, w, vzi, vzi);
      ch |= 4u << r;
    }
    if (lane == r) {
      eS[p] = eSp; fS[p] = fSp;
      eS[q] = eSq; fS[q] = fSq;
    }
    wG = xmax(wG, o.tG);
    wcnt += (unsigned long long)o.trans;
  };
  apply(x0, y0, v0, w0, 0, p0, q0);
  if (two && !werr)
    apply(x1, y1, v1, w1, 1, p1, q1);
  return ch;
}

// Stores the changed columns of a unit after its round and records the step
// of each G column's last change; ch0 and ch1 are the change bits (pair2's
// form) of its two steps.  Returns the elements moved (loads included).
template <typename T, int E>
__device__ __forceinline__ unsigned long long unit_store(T *Gr, T *Vr, const int m, const int n, const int nE,
                                                         const int lane, const Step<T> &S, int *mod, const int g0,
                                                         const unsigned I, const unsigned J, const unsigned ch0,
                                                         const unsigned ch1, T (&ga)[E], T (&gb)[E], T (&gc)[E],
                                                         T (&gd)[E], T (&va)[2], T (&vb)[2], T (&vc)[2], T (&vd)[2])
{
  const bool diag = (I == J);
  const int g1 = g0 + 1;
  const unsigned cA = 2u * I, cB = 2u * I + 1u, cC = 2u * J, cD = 2u * J + 1u;
  // changes per register column: a by pair 0 at both steps, b by pair 1
  // at both, c by pair 0 then 1, d by pair 1 then 0 (for a block met by
  // itself d holds column 2I+1, changed by pair 0 at both steps)
  const unsigned s0a = ch0 & 5u, s1a = ch1 & 5u, s0b = (ch0 >> 1) & 5u, s1b = (ch1 >> 1) & 5u;
  const unsigned chA[4] = {s0a | s1a, s0b | s1b, s0a | s1b, diag ? (s0a | s1a) : (s0b | s1a)};
  const bool last1[4] = {(s1a & 1u) != 0u, (s1b & 1u) != 0u, (s1b & 1u) != 0u, (s1a & 1u) != 0u};
  const unsigned col[4] = {cA, cB, cC, diag ? cB : cD};
  unsigned long long xf = (unsigned long long)(diag ? 2 : 4) * (unsigned long long)(m + n);
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    if (diag && (k == 1 || k == 2))
      continue;
    T (&g)[E] = (k == 0) ? ga : (k == 1) ? gb : (k == 2) ? gc : gd;
    T (&v)[2] = (k == 0) ? va : (k == 1) ? vb : (k == 2) ? vc : vd;
    if ((chA[k] & 1u) || S.sG) {
      store_col<T, E>(Gr, m, nE, col[k], lane, g);
      if (lane == 0)
        mod[col[k]] = last1[k] ? g1 : g0;
      xf += (unsigned long long)m;
    }
    if ((chA[k] & 4u) || S.sV) {
      store_col<T, 2>(Vr, n, 2, col[k], lane, v);
      xf += (unsigned long long)n;
    }
  }
  return xf;
}

// Register footprint of the block form: the next unit is prefetched while
// its four G columns take at most 8 registers each.
template <typename T, int E> struct BlockShape {
  static constexpr bool PF = (E * (int)(sizeof(T) / 4)) <= 8;
};

// Block round s (steps 2s and 2s + 1, numbered g0 and g0 + 1 in the run) for
// the units of one wave (units wave, wave + W, ...).  A unit none of whose
// columns changed since step g0 - n (one sweep before) is skipped: each of
// its pairs would repeat its no-op evaluation of the previous sweep.
template <typename T, int E, int W, bool PF>
__device__ __forceinline__ void block_round(T *Gr, T *Vr, T *eS, T *fS, const int m, const int n, const int nE,
                                            const unsigned s, const int wave, const int lane, const Step<T> &S,
                                            T &wG, unsigned long long &wcnt, int &werr, int *mod, const int g0,
                                            unsigned long long &wev, unsigned long long &wxf)
{
  const unsigned nb = (unsigned)n >> 1u;
  const int nu = (int)br_units(nb, s);
  if (wave >= nu)
    return;
  const int cnt0 = (nu - wave + W - 1) / W;
  int cnt = cnt0, slot = lane;
  if ((S.sG == 0) && (S.sV == 0) && (g0 > n)) {
    bool need = false;
    if (lane < cnt0) {
      unsigned I, J;
      br_unit(nb, s, (unsigned)(wave + W * lane), I, J);
      const int lim = g0 - n;
      need = (mod[2u * I] >= lim) || (mod[2u * I + 1u] >= lim) || (mod[2u * J] >= lim) || (mod[2u * J + 1u] >= lim);
    }
    const unsigned long long mk = __ballot(need);
    cnt = __popcll(mk);
    if (cnt == 0)
      return;
    const int below = __popcll(mk & ((1ull << lane) - 1ull));
    const int pos = need ? below : (cnt + lane - below);
    slot = __builtin_amdgcn_ds_permute(pos << 2, lane);
  }
  auto unit_of = [&](const int i, unsigned &I, unsigned &J) {
    const int k = __builtin_amdgcn_readlane(slot, (i < cnt) ? i : (cnt - 1));
    br_unit(nb, s, (unsigned)(wave + W * k), I, J);
  };
  T zi[E], vzi[2];
  auto load_unit = [&](const unsigned I, const unsigned J, T (&a)[E], T (&b)[E], T (&c)[E], T (&d)[E], T (&va)[2],
                       T (&vb)[2], T (&vc)[2], T (&vd)[2]) {
    load_pair<T, false, E>(Gr, nullptr, m, nE, (int)(2u * I), (int)(2u * I + 1u), lane, a, b, zi, zi);
    load_pair<T, false, E>(Gr, nullptr, m, nE, (int)(2u * J), (int)(2u * J + 1u), lane, c, d, zi, zi);
    load_vpair<T, false>(Vr, nullptr, n, (int)(2u * I), (int)(2u * I + 1u), lane, va, vb, vzi, vzi);
    load_vpair<T, false>(Vr, nullptr, n, (int)(2u * J), (int)(2u * J + 1u), lane, vc, vd, vzi, vzi);
  };
  T ga[E], gb[E], gc[E], gd[E], va[2], vb[2], vc[2], vd[2];
  unsigned I = 0u, J = 0u;
  if constexpr (PF) {
    unit_of(0, I, J);
    load_unit(I, J, ga, gb, gc, gd, va, vb, vc, vd);
  }
  for (int i = 0; i < cnt; ++i) {
    unsigned In = 0u, Jn = 0u;
    T na[E], nbb[E], nc[E], nd[E], nva[2], nvb[2], nvc[2], nvd[2];
    if constexpr (PF) {
      unit_of((i + 1 < cnt) ? (i + 1) : i, In, Jn);
      load_unit(In, Jn, na, nbb, nc, nd, nva, nvb, nvc, nvd);
    }
    else {
      unit_of(i, I, J);
      load_unit(I, J, ga, gb, gc, gd, va, vb, vc, vd);
    }
    ready_pair<T, false, E>(m, nE, S.sG, lane, ga, gb, zi, zi);
    ready_pair<T, false, E>(m, nE, S.sG, lane, gc, gd, zi, zi);
    if (S.sV) {
      scale_vpair<T>(S.sV, va, vb, vzi, vzi);
      scale_vpair<T>(S.sV, vc, vd, vzi, vzi);
    }
    // a block met by itself runs its pair (2I, 2I+1) at both steps as pair
    // 0: column 2I+1 moves to register column c for the first step and on
    // to d for the second
    const bool diag = (I == J);
    if (diag) {
#pragma unroll
      for (int j = 0; j < E; ++j)
        gc[j] = gb[j];
      vc[0] = vb[0];
      vc[1] = vb[1];
    }
    const unsigned cA = 2u * I, cB = 2u * I + 1u, cC = 2u * J, cD = 2u * J + 1u;
    const unsigned ch0 = pair2<T, E>(ga, gc, gb, gd, va, vc, vb, vd, !diag, cA, diag ? cB : cC, cB, cD, nE, lane, S,
                                     eS, fS, wG, wcnt, werr);
    if (werr)
      return;
    if (diag) {
#pragma unroll
      for (int j = 0; j < E; ++j)
        gd[j] = gc[j];
      vd[0] = vc[0];
      vd[1] = vc[1];
    }
    const unsigned ch1 = pair2<T, E>(ga, gd, gb, gc, va, vd, vb, vc, !diag, cA, diag ? cB : cD, cB, cC, nE, lane, S,
                                     eS, fS, wG, wcnt, werr);
    if (werr)
      return;
    wev += diag ? 2ull : 4ull;
    wxf += unit_store<T, E>(Gr, Vr, m, n, nE, lane, S, mod, g0, I, J, ch0, ch1, ga, gb, gc, gd, va, vb, vc, vd);
    if constexpr (PF) {
      I = In;
      J = Jn;
#pragma unroll
      for (int j = 0; j < E; ++j) {
        ga[j] = na[j]; gb[j] = nbb[j]; gc[j] = nc[j]; gd[j] = nd[j];
      }
      va[0] = nva[0]; va[1] = nva[1]; vb[0] = nvb[0]; vb[1] = nvb[1];
      vc[0] = nvc[0]; vc[1] = nvc[1]; vd[0] = nvd[0]; vd[1] = nvd[1];
    }
  }
}

// Four pairs of one step (two units of a wave), reduced and decided
// together: lane r & 3 decides pair r, so that one decision stream serves
// four pairs; pair r is applied when bit r of `valid` is set.  Returns bit r
// when pair r changed its G columns and bit 4 + r when it changed its V
// columns; werr as in pair2.
template <typename T, int E>
__device__ __forceinline__ unsigned pair4(T (&x0)[E], T (&y0)[E], T (&x1)[E], T (&y1)[E], T (&x2)[E], T (&y2)[E],
                                          T (&x3)[E], T (&y3)[E], T (&v0)[2], T (&w0)[2], T (&v1)[2], T (&w1)[2],
                                          T (&v2)[2], T (&w2)[2], T (&v3)[2], T (&w3)[2], const unsigned valid,
                                          const unsigned (&pq)[8], const int nE, const int lane, const Step<T> &S,
                                          T *eS, T *fS, T &wG, unsigned long long &wcnt, int &werr)
{
  typedef typename Acc<T>::A A;
  constexpr int B = 64;
  T zi[E], vzi[2]; // imaginary parts of real data
#pragma unroll
  for (int j = 0; j < E; ++j)
    zi[j] = (T)0;
  vzi[0] = vzi[1] = (T)0;
  T mx0, my0, mx1, my1, mx2, my2, mx3, my3;
  pair_maxima<T, false, E>(x0, y0, zi, zi, nE, mx0, my0);
  pair_maxima<T, false, E>(x1, y1, zi, zi, nE, mx1, my1);
  pair_maxima<T, false, E>(x2, y2, zi, zi, nE, mx2, my2);
  pair_maxima<T, false, E>(x3, y3, zi, zi, nE, mx3, my3);
  red_max4(mx0, my0, mx1, my1);
  red_max4(mx2, my2, mx3, my3);
  const int kx0 = max_exp(mx0), ky0 = max_exp(my0), kx1 = max_exp(mx1), ky1 = max_exp(my1);
  const int kx2 = max_exp(mx2), ky2 = max_exp(my2), kx3 = max_exp(mx3), ky3 = max_exp(my3);
  DD<A> sx0{0, 0}, sy0{0, 0}, dr0{0, 0}, sx1{0, 0}, sy1{0, 0}, dr1{0, 0};
  DD<A> sx2{0, 0}, sy2{0, 0}, dr2{0, 0}, sx3{0, 0}, sy3{0, 0}, dr3{0, 0}, di{0, 0};
  pair_sums<T, false, E>(x0, y0, zi, zi, nE, kx0, ky0, sx0, sy0, dr0, di);
  pair_sums<T, false, E>(x1, y1, zi, zi, nE, kx1, ky1, sx1, sy1, dr1, di);
  red_dd4<A, B>(sx0, sy0, dr0, sx1, nullptr);
  pair_sums<T, false, E>(x2, y2, zi, zi, nE, kx2, ky2, sx2, sy2, dr2, di);
  red_dd4<A, B>(sy1, dr1, sx2, sy2, nullptr);
  pair_sums<T, false, E>(x3, y3, zi, zi, nE, kx3, ky3, sx3, sy3, dr3, di);
  red_dd4<A, B>(dr2, sx3, sy3, dr3, nullptr);
  const int rl = lane & 3;
  const bool b0 = (rl & 1) != 0, b1 = (rl & 2) != 0;
  auto sel = [&](auto a0, auto a1, auto a2, auto a3) { return b1 ? (b0 ? a3 : a2) : (b0 ? a1 : a0); };
  auto seld = [&](DD<A> a0, DD<A> a1, DD<A> a2, DD<A> a3) {
    return dd_sel(b1, dd_sel(b0, a3, a2), dd_sel(b0, a1, a0));
  };
  const DD<A> dz{0, 0};
  const PairScal<T> Pl = pair_scal<T, false>(sel(mx0, mx1, mx2, mx3), sel(my0, my1, my2, my3),
                                             sel(kx0, kx1, kx2, kx3), sel(ky0, ky1, ky2, ky3),
                                             seld(sx0, sx1, sx2, sx3), seld(sy0, sy1, sy2, sy3),
                                             seld(dr0, dr1, dr2, dr3), dz);
  const PairDec<T> Dl = pair_decide<T, false>(Pl, S.tol, S.gst);
  unsigned ch = 0u;
  auto apply = [&](T (&x)[E], T (&y)[E], T (&v)[2], T (&w)[2], const int r) {
    const PairDec<T> D = lane_dec(Dl, r);
    const bool gsr = (D.code == 3) || (D.code == -3);
    const PairScal<T> P = gsr ? lane_scal(Pl, r) : Pl;
    T eSp, fSp, eSq, fSq;
    const PairOut<T, false> o = pair_apply<T, false, B, E, false>(0, x, y, zi, zi, nullptr, nullptr, nullptr, nullptr,
                                                                  P, D, 0, eSp, fSp, eSq, fSq, nullptr, lane, nE);
    int err = D.err;
    if (!err && (__ballot(!isfinite_(o.tG)) != 0ull))
      err = -580;
    if (err) {
      werr = err;
      return;
    }
    if (D.code != 1)
      ch |= 1u << r;
    if ((D.code != 1) && (D.code != 3)) {
      v_update<T, false>(D, v, w, vzi, vzi);
      ch |= 16u << r;
    }
    if (lane == r) {
      eS[pq[2 * r]] = eSp; fS[pq[2 * r]] = fSp;
      eS[pq[2 * r + 1]] = eSq; fS[pq[2 * r + 1]] = fSq;
    }
    wG = xmax(wG, o.tG);
    wcnt += (unsigned long long)o.trans;
  };
  if (valid & 1u)
    apply(x0, y0, v0, w0, 0);
  if ((valid & 2u) && !werr)
    apply(x1, y1, v1, w1, 1);
  if ((valid & 4u) && !werr)
    apply(x2, y2, v2, w2, 2);
  if ((valid & 8u) && !werr)
    apply(x3, y3, v3, w3, 3);
  return ch;
}

// Block round with two units at a time per wave (short columns, up to 4
// row chunks): the four pairs of a step share one decision stream (pair4),
// so that a pair's decision costs a quarter of a pass instead of a half.
// No prefetch (two units' columns already take the registers of one
// prefetched unit); an odd last unit runs with an unapplied copy of itself.
template <typename T, int E, int W>
__device__ __forceinline__ void block_round2(T *Gr, T *Vr, T *eS, T *fS, const int m, const int n, const int nE,
                                             const unsigned s, const int wave, const int lane, const Step<T> &S,
                                             T &wG, unsigned long long &wcnt, int &werr, int *mod, const int g0,
                                             unsigned long long &wev, unsigned long long &wxf)
{
  const unsigned nb = (unsigned)n >> 1u;
  const int nu = (int)br_units(nb, s);
  if (wave >= nu)
    return;
  const int cnt0 = (nu - wave + W - 1) / W;
  int cnt = cnt0, slot = lane;
  if ((S.sG == 0) && (S.sV == 0) && (g0 > n)) {
    bool need = false;
    if (lane < cnt0) {
      unsigned I, J;
      br_unit(nb, s, (unsigned)(wave + W * lane), I, J);
      const int lim = g0 - n;
      need = (mod[2u * I] >= lim) || (mod[2u * I + 1u] >= lim) || (mod[2u * J] >= lim) || (mod[2u * J + 1u] >= lim);
    }
    const unsigned long long mk = __ballot(need);
    cnt = __popcll(mk);
    if (cnt == 0)
      return;
    const int below = __popcll(mk & ((1ull << lane) - 1ull));
    const int pos = need ? below : (cnt + lane - below);
    slot = __builtin_amdgcn_ds_permute(pos << 2, lane);
  }
  auto unit_of = [&](const int i, unsigned &I, unsigned &J) {
    const int k = __builtin_amdgcn_readlane(slot, (i < cnt) ? i : (cnt - 1));
    br_unit(nb, s, (unsigned)(wave + W * k), I, J);
  };
  T zi[E], vzi[2];
  auto load_unit = [&](const unsigned I, const unsigned J, T (&a)[E], T (&b)[E], T (&c)[E], T (&d)[E], T (&va)[2],
                       T (&vb)[2], T (&vc)[2], T (&vd)[2]) {
    load_pair<T, false, E>(Gr, nullptr, m, nE, (int)(2u * I), (int)(2u * I + 1u), lane, a, b, zi, zi);
    load_pair<T, false, E>(Gr, nullptr, m, nE, (int)(2u * J), (int)(2u * J + 1u), lane, c, d, zi, zi);
    load_vpair<T, false>(Vr, nullptr, n, (int)(2u * I), (int)(2u * I + 1u), lane, va, vb, vzi, vzi);
    load_vpair<T, false>(Vr, nullptr, n, (int)(2u * J), (int)(2u * J + 1u), lane, vc, vd, vzi, vzi);
    ready_pair<T, false, E>(m, nE, S.sG, lane, a, b, zi, zi);
    ready_pair<T, false, E>(m, nE, S.sG, lane, c, d, zi, zi);
    if (S.sV) {
      scale_vpair<T>(S.sV, va, vb, vzi, vzi);
      scale_vpair<T>(S.sV, vc, vd, vzi, vzi);
    }
  };
  for (int i = 0; i < cnt; i += 2) {
    const bool hasB = (i + 1 < cnt);
    unsigned Ia, Ja, Ib, Jb;
    unit_of(i, Ia, Ja);
    unit_of(i + 1, Ib, Jb); // the last unit again when there is no second one
    T ga[E], gb[E], gc[E], gd[E], va[2], vb[2], vc[2], vd[2];
    T ha[E], hb[E], hc[E], hd[E], ua[2], ub[2], uc[2], ud[2];
    load_unit(Ia, Ja, ga, gb, gc, gd, va, vb, vc, vd);
    load_unit(Ib, Jb, ha, hb, hc, hd, ua, ub, uc, ud);
    const bool dA = (Ia == Ja), dB = (Ib == Jb);
    if (dA) {
#pragma unroll
      for (int j = 0; j < E; ++j)
        gc[j] = gb[j];
      vc[0] = vb[0];
      vc[1] = vb[1];
    }
    if (dB) {
#pragma unroll
      for (int j = 0; j < E; ++j)
        hc[j] = hb[j];
      uc[0] = ub[0];
      uc[1] = ub[1];
    }
    const unsigned valid = 1u | (dA ? 0u : 2u) | (hasB ? (4u | (dB ? 0u : 8u)) : 0u);
    const unsigned aA = 2u * Ia, bA = aA + 1u, cA = 2u * Ja, dA_ = cA + 1u;
    const unsigned aB = 2u * Ib, bB = aB + 1u, cB = 2u * Jb, dB_ = cB + 1u;
    const unsigned pq0[8] = {aA, dA ? bA : cA, bA, dA_, aB, dB ? bB : cB, bB, dB_};
    const unsigned c0 = pair4<T, E>(ga, gc, gb, gd, ha, hc, hb, hd, va, vc, vb, vd, ua, uc, ub, ud, valid, pq0, nE, lane,
                                    S, eS, fS, wG, wcnt, werr);
    if (werr)
      return;
    if (dA) {
#pragma unroll
      for (int j = 0; j < E; ++j)
        gd[j] = gc[j];
      vd[0] = vc[0];
      vd[1] = vc[1];
    }
    if (dB) {
#pragma unroll
      for (int j = 0; j < E; ++j)
        hd[j] = hc[j];
      ud[0] = uc[0];
      ud[1] = uc[1];
    }
    const unsigned pq1[8] = {aA, dA ? bA : dA_, bA, cA, aB, dB ? bB : dB_, bB, cB};
    const unsigned c1 = pair4<T, E>(ga, gd, gb, gc, ha, hd, hb, hc, va, vd, vb, vc, ua, ud, ub, uc, valid, pq1, nE, lane,
                                    S, eS, fS, wG, wcnt, werr);
    if (werr)
      return;
    // the pair2 form of the change bits of each unit: G bits r, V bits 2 + r
    const unsigned a0 = (c0 & 3u) | (((c0 >> 4) & 3u) << 2), a1 = (c1 & 3u) | (((c1 >> 4) & 3u) << 2);
    const unsigned b0 = ((c0 >> 2) & 3u) | (((c0 >> 6) & 3u) << 2), b1 = ((c1 >> 2) & 3u) | (((c1 >> 6) & 3u) << 2);
    wev += dA ? 2ull : 4ull;
    wxf += unit_store<T, E>(Gr, Vr, m, n, nE, lane, S, mod, g0, Ia, Ja, a0, a1, ga, gb, gc, gd, va, vb, vc, vd);
    if (hasB) {
      wev += dB ? 2ull : 4ull;
      wxf += unit_store<T, E>(Gr, Vr, m, n, nE, lane, S, mod, g0, Ib, Jb, b0, b1, ha, hb, hc, hd, ua, ub, uc, ud);
    }
  }
}

// ------------------------------------------------------- lane-group steps
// A wave works on P = 64 / LP pairs at once, one per group of LP lanes
// (LP = 16: a DPP row; 32: two rows; 64: the whole wave).  Lane t of a group
// holds rows t + LP j (j < E) of the pair's G columns and of its V columns
// (n <= m rows, so E chunks cover them too).  The reductions run inside the
// groups (DPP within a row, then permlane16 / permlane32 swaps) and leave
// the group totals in every lane of the group, so every lane decides its own
// group's pair (the decisions differ between groups, not within one) and
// the update needs no broadcast.  A permuting update (codes -1, -2) swaps
// the destination columns instead of the registers.
template <typename T, int LP> __device__ __forceinline__ T group_max(T v)
{
  v = row_max<T, (LP < 16) ? LP : 16>(v);
  if (LP >= 32) {
    T a = v, b = v;
    swap_<false>(a, b); // a: rows (v0, v0, v2, v2), b: rows (v1, v1, v3, v3)
    v = xmax(a, b);
  }
  if (LP >= 64) {
    T a = v, b = v;
    swap_<true>(a, b);
    v = xmax(a, b);
  }
  return v;
}
template <typename A, int LP> __device__ __forceinline__ DD<A> group_sum(DD<A> v)
{
  v = row_sum<A, (LP < 16) ? LP : 16>(v);
  if (LP >= 32) {
    DD<A> a = v, b = v;
    swap_<false>(a.s, b.s);
    swap_<false>(a.c, b.c);
    v = dd_add(a, b); // dd_add is commutative bit for bit: every lane agrees
  }
  if (LP >= 64) {
    DD<A> a = v, b = v;
    swap_<true>(a.s, b.s);
    swap_<true>(a.c, b.c);
    v = dd_add(a, b);
  }
  return v;
}

// Columns p, q of a rows-row matrix (ld rows) into registers, row t + LP j;
// all E chunks are loaded (rows past the end read the last row), so the
// number of loads in flight is fixed.
template <typename T, bool CPLX, int E>
__device__ __forceinline__ void load_cols(const T *Ar, const T *Ai, const int rows, const unsigned p, const unsigned q,
                                          const int t, const int LP, T (&x)[E], T (&y)[E], T (&xi)[E], T (&yi)[E])
{
  const long long op = (long long)p * rows, oq = (long long)q * rows;
#pragma unroll
  for (int j = 0; j < E; ++j) {
    const int i = t + LP * j;
    const int ic = (i < rows) ? i : (rows - 1);
    x[j] = Ar[op + ic];
    y[j] = Ar[oq + ic];
    if constexpr (CPLX) {
      xi[j] = Ai[op + ic];
      yi[j] = Ai[oq + ic];
    }
    else
      xi[j] = yi[j] = (T)0;
  }
}

template <typename T, bool CPLX, int E>
__device__ __forceinline__ void store_cols(T *Ar, T *Ai, const int rows, const int nE, const unsigned p, const unsigned q,
                                           const int t, const int LP, const T (&x)[E], const T (&y)[E], const T (&xi)[E],
                                           const T (&yi)[E])
{
  const long long op = (long long)p * rows, oq = (long long)q * rows;
#pragma unroll
  for (int j = 0; j < E; ++j) {
    if (j >= nE)
      break;
    const int i = t + LP * j;
    if (i < rows) {
      Ar[op + i] = x[j];
      Ar[oq + i] = y[j];
      if constexpr (CPLX) {
        Ai[op + i] = xi[j];
        Ai[oq + i] = yi[j];
      }
    }
  }
}

// Real rotation (x, y) := ((x + t y) c, (y - t x) c) of E chunks, the
// scaling skipped where c == 1 (the product by one is exact).
template <typename T, bool CPLX, int E>
__device__ __forceinline__ void grp_rot_real(T (&x)[E], T (&y)[E], T (&xi)[E], T (&yi)[E], const T c, const T at,
                                             const bool c1, const int nE)
{
#pragma unroll
  for (int j = 0; j < E; ++j) {
    if (j >= nE)
      break;
    const T u = fma_(y[j], at, x[j]), v = fma_(-x[j], at, y[j]);
    x[j] = u;
    y[j] = v;
    if constexpr (CPLX) {
      const T u2 = fma_(yi[j], at, xi[j]), v2 = fma_(-xi[j], at, yi[j]);
      xi[j] = u2;
      yi[j] = v2;
    }
  }
  if (!c1) {
#pragma unroll
    for (int j = 0; j < E; ++j) {
      if (j >= nE)
        break;
      x[j] = x[j] * c;
      y[j] = y[j] * c;
      if constexpr (CPLX) {
        xi[j] = xi[j] * c;
        yi[j] = yi[j] * c;
      }
    }
  }
}

// The lane's max |element| of a pair after grp_rot_real (one product per
// element, so the IEEE maximum).
template <typename T, bool CPLX, int E>
__device__ __forceinline__ T grp_max_real(const T (&x)[E], const T (&y)[E], const T (&xi)[E], const T (&yi)[E],
                                          const int nE)
{
  T tG = (T)0;
#pragma unroll
  for (int j = 0; j < E; ++j) {
    if (j >= nE)
      break;
    tG = max1p(tG, max1p(vabs(x[j]), vabs(y[j])));
    if constexpr (CPLX)
      tG = max1p(tG, max1p(vabs(xi[j]), vabs(yi[j])));
  }
  return tG;
}

// Complex rotation with a complex tangent at + i sat (zjrot_, zjrot.c:36-60).
template <typename T, int E>
__device__ __forceinline__ void grp_rot_cplx(T (&x)[E], T (&y)[E], T (&xi)[E], T (&yi)[E], const T c, const T at,
                                             const T sat, const bool c1, const int nE)
{
  const T cna = -at;
#pragma unroll
  for (int j = 0; j < E; ++j) {
    if (j >= nE)
      break;
    const T x_r = x[j], x_i = xi[j], y_r = y[j], y_i = yi[j];
    T ar = fma_(y_r, at, fma_(-y_i, sat, x_r));
    T ai = fma_(y_r, sat, fma_(y_i, at, x_i));
    T br = fma_(x_r, cna, fma_(-x_i, sat, y_r));
    T bi = fma_(x_r, sat, fma_(x_i, cna, y_i));
    if (!c1) { ar = ar * c; ai = ai * c; br = br * c; bi = bi * c; }
    x[j] = ar; xi[j] = ai; y[j] = br; yi[j] = bi;
  }
}

// The lane's max |element| of a pair after grp_rot_cplx (two products
// summed per element: the reference's ordered maximum, in its order).
template <typename T, int E>
__device__ __forceinline__ T grp_max_cplx(const T (&x)[E], const T (&y)[E], const T (&xi)[E], const T (&yi)[E],
                                          const int nE)
{
  T tG = (T)0;
#pragma unroll
  for (int j = 0; j < E; ++j) {
    if (j >= nE)
      break;
    tG = xmax(tG, xmax(xmax(vabs(x[j]), vabs(xi[j])), xmax(vabs(y[j]), vabs(yi[j]))));
  }
  return tG;
}

// Gram-Schmidt correction (dgsscl_/zgsscl_, dgsscl.c:5-85, zgsscl.c:5-119)
// of pair_apply, per lane; for code -3 the result lands in x's place and y
// takes the old x... as there: x := y, y := the corrected column.
template <typename T, bool CPLX, int E>
__device__ __forceinline__ T grp_gs(T (&xr)[E], T (&yr)[E], T (&xi)[E], T (&yi)[E], const PairScal<T> &P,
                                    const bool perm, const int nE)
{
  const T ex = P.ex, fx = P.fx, ey = P.ey, fy = P.fy;
  const T x_e = -ex, y_e = -ey;
  T tG = (T)0;
  if (!perm) {
    const T fy_x = fy / fx;
    const T tfr = -P.a21r * fy_x;
    const T tfi = P.a21i * fy_x;
#pragma unroll
    for (int j = 0; j < E; ++j) {
      if (j >= nE)
        break;
      const T xs = scalef_fin(xr[j], x_e);
      const T ys = scalef_fin(yr[j], y_e);
      if constexpr (CPLX) {
        const T xsi = scalef_fin(xi[j], x_e), ysi = scalef_fin(yi[j], y_e);
        const T nr = fma_(tfr, xs, fma_(-tfi, xsi, ys));
        const T ni = fma_(tfr, xsi, fma_(tfi, xs, ysi));
        yr[j] = scalef_fin(nr, ey);
        yi[j] = scalef_fin(ni, ey);
        tG = xmax(tG, xmax(vabs(yr[j]), vabs(yi[j])));
      }
      else {
        yr[j] = scalef_fin(fma_(tfr, xs, ys), ey);
        tG = max1p(tG, vabs(yr[j]));
      }
    }
  }
  else {
    const T fx_y = fx / fy;
    const T tfr = -P.a21r * fx_y;
    const T tfi = -P.a21i * fx_y;
#pragma unroll
    for (int j = 0; j < E; ++j) {
      if (j >= nE)
        break;
      const T xs = scalef_fin(xr[j], x_e);
      const T ys = scalef_fin(yr[j], y_e);
      if constexpr (CPLX) {
        const T xsi = scalef_fin(xi[j], x_e), ysi = scalef_fin(yi[j], y_e);
        const T nr = fma_(tfr, ys, fma_(-tfi, ysi, xs));
        const T ni = fma_(tfr, ysi, fma_(tfi, ys, xsi));
        xr[j] = yr[j];
        xi[j] = yi[j];
        yr[j] = scalef_fin(nr, ex);
        yi[j] = scalef_fin(ni, ex);
        tG = xmax(tG, xmax(vabs(yr[j]), vabs(yi[j])));
      }
      else {
        xr[j] = yr[j];
        yr[j] = scalef_fin(fma_(tfr, ys, xs), ex);
        tG = max1p(tG, vabs(yr[j]));
      }
    }
  }
  return tG;
}

// Whether the lane-group form prefetches the next group's pairs.  It does
// not: the double buffer takes the FP64 short-column kernel from 144 to 180
// registers, i.e. from three waves per SIMD to two, and a wave of a matrix
// narrower than 8W columns has a single group iteration, whose "next" pairs
// would be the same columns loaded again.  Without it 4096 matrices of 64x32
// take 11.9 ms instead of 13.1 ms, and 16384 of n <= 128 are not slower.
template <typename T, bool CPLX, int E> struct GroupShape {
  static constexpr bool PF = false;
};

// One step of the strategy for the pairs of one wave (pairs wave, wave + W,
// wave + 2W, ... of step st of ordering ord, pjs id 4 or 6), P pairs at a
// time in lane groups: load (the next P prefetched when PF), reduce, decide
// per lane, update in registers, store.  Accumulates this wave's per-lane
// max of the updated G elements, its transformation count and its first
// error.
template <typename T, bool CPLX, int E, int W, bool PF, int LG>
__device__ __forceinline__ void group_step(T *Gr, T *Gi, T *Vr, T *Vi, T *eS, T *fS, const int m, const int n,
                                           const int ord, const unsigned st, const int wave, const int lane,
                                           const Step<T> &S, T &wG, T &wV, unsigned long long &wcnt, int &werr,
                                           unsigned long long &wev, unsigned long long *wxf, const bool gmem)
{
  typedef typename Acc<T>::A A;
  constexpr int lg = LG, LP = 1 << LG, P = 64 >> LG; // lanes per group, groups per wave
  // (wxf counts the elements moved through device memory: G's columns only
  // when G is there, gmem, V's always)
  constexpr unsigned long long PL = CPLX ? 2ull : 1ull; // planes
  const int npairs = n >> 1;
  if (wave >= npairs)
    return;
  const int cnt = (npairs - wave + W - 1) / W; // pairs of this wave
  wev += (unsigned long long)cnt;
  if (wxf)
    *wxf += (unsigned long long)cnt * 2ull * PL * (unsigned long long)((gmem ? m : 0) + n);
  const int nit = (cnt + P - 1) >> (6 - lg);
  const int g = lane >> lg, t = lane & (LP - 1);
  const int nE = (m + LP - 1) >> lg, nEV = (n + LP - 1) >> lg;
  // pair of this group in iteration it (past the end: the wave's last pair
  // again, computed and not stored)
  auto pair_of = [&](const int it, unsigned &pp, unsigned &qq) -> bool {
    int i = it * P + g;
    const bool v = (i < cnt);
    i = v ? i : (cnt - 1);
    step_pair(ord, (unsigned)n, st, (unsigned)(wave + W * i), pp, qq);
    return v;
  };
  T xr[E], yr[E], xi[E], yi[E], vx[E], vy[E], vxi[E], vyi[E];
  unsigned p, q;
  bool valid;
  if constexpr (PF) {
    valid = pair_of(0, p, q);
    load_cols<T, CPLX, E>(Vr, Vi, n, p, q, t, LP, vx, vy, vxi, vyi);
    load_cols<T, CPLX, E>(Gr, Gi, m, p, q, t, LP, xr, yr, xi, yi);
  }
  for (int it = 0; it < nit; ++it) {
    unsigned pn, qn;
    bool validn = false;
    T nxr[E], nyr[E], nxi[E], nyi[E], nvx[E], nvy[E], nvxi[E], nvyi[E];
    if constexpr (PF) {
      validn = pair_of((it + 1 < nit) ? (it + 1) : it, pn, qn);
      load_cols<T, CPLX, E>(Vr, Vi, n, pn, qn, t, LP, nvx, nvy, nvxi, nvyi);
      load_cols<T, CPLX, E>(Gr, Gi, m, pn, qn, t, LP, nxr, nyr, nxi, nyi);
    }
    else {
      valid = pair_of(it, p, q);
      load_cols<T, CPLX, E>(Vr, Vi, n, p, q, t, LP, vx, vy, vxi, vyi);
      load_cols<T, CPLX, E>(Gr, Gi, m, p, q, t, LP, xr, yr, xi, yi);
    }
    // ---- rows past m to zero, pending rescalings
#pragma unroll
    for (int j = 0; j < E; ++j) {
      if (j >= nE)
        break;
      const bool in = (t + LP * j) < m;
      xr[j] = in ? xr[j] : (T)0;
      yr[j] = in ? yr[j] : (T)0;
      if constexpr (CPLX) {
        xi[j] = in ? xi[j] : (T)0;
        yi[j] = in ? yi[j] : (T)0;
      }
      if (S.sG) {
        xr[j] = ldexp_(xr[j], S.sG);
        yr[j] = ldexp_(yr[j], S.sG);
        if constexpr (CPLX) {
          xi[j] = ldexp_(xi[j], S.sG);
          yi[j] = ldexp_(yi[j], S.sG);
        }
      }
    }
    if (S.sV) {
#pragma unroll
      for (int j = 0; j < E; ++j) {
        vx[j] = ldexp_(vx[j], S.sV);
        vy[j] = ldexp_(vy[j], S.sV);
        vxi[j] = ldexp_(vxi[j], S.sV);
        vyi[j] = ldexp_(vyi[j], S.sV);
      }
    }
    // ---- reduce within the groups
    T mx, my;
    pair_maxima<T, CPLX, E>(xr, yr, xi, yi, nE, mx, my);
    mx = group_max<T, LP>(mx);
    my = group_max<T, LP>(my);
    const int kx = max_exp(mx), ky = max_exp(my);
    DD<A> sx{0, 0}, sy{0, 0}, dr{0, 0}, di{0, 0};
    pair_sums<T, CPLX, E>(xr, yr, xi, yi, nE, kx, ky, sx, sy, dr, di);
    sx = group_sum<A, LP>(sx);
    sy = group_sum<A, LP>(sy);
    dr = group_sum<A, LP>(dr);
    if constexpr (CPLX)
      di = group_sum<A, LP>(di);
    // ---- every lane decides its group's pair
    const PairScal<T> Pd = pair_scal<T, CPLX>(mx, my, kx, ky, sx, sy, dr, di);
    const PairDec<T> D = pair_decide<T, CPLX>(Pd, S.tol, S.gst);
    // ---- update in registers
    const int code = D.code;
    const bool ok = (D.err == 0);
    const bool two = ok && ((code == 2) || (code == -2));
    const bool at0 = (D.at == (T)0) && (!CPLX || (D.sat == (T)0));
    const bool rotr = two && !at0 && (!CPLX || (D.sat == (T)0));
    const bool rotc = CPLX && two && (D.sat != (T)0);
    const bool gs = ok && ((code == 3) || (code == -3));
    const bool c1 = (D.c == (T)1);
    // destination columns exchanged: a swap (-1), or a rotation or a unit
    // swap with the permutation (-2)
    const bool dsw = ok && ((code == -1) || ((code == -2) && (!at0 || c1)));
    T tG = (T)0, tV = (T)0;
    if (rotr) {
      grp_rot_real<T, CPLX, E>(xr, yr, xi, yi, D.c, D.at, c1, nE);
      tG = grp_max_real<T, CPLX, E>(xr, yr, xi, yi, nE);
      grp_rot_real<T, CPLX, E>(vx, vy, vxi, vyi, D.c, D.at, c1, nEV);
      if (S.vtrack)
        tV = grp_max_real<T, CPLX, E>(vx, vy, vxi, vyi, nEV);
    }
    if constexpr (CPLX) {
      if (rotc) {
        grp_rot_cplx<T, E>(xr, yr, xi, yi, D.c, D.at, D.sat, c1, nE);
        tG = grp_max_cplx<T, E>(xr, yr, xi, yi, nE);
        grp_rot_cplx<T, E>(vx, vy, vxi, vyi, D.c, D.at, D.sat, c1, nEV);
        if (S.vtrack)
          tV = grp_max_cplx<T, E>(vx, vy, vxi, vyi, nEV);
      }
    }
    if (gs) {
      const bool perm = (code < 0);
      tG = grp_gs<T, CPLX, E>(xr, yr, xi, yi, Pd, perm, nE);
      if (perm) {
#pragma unroll
        for (int j = 0; j < E; ++j) {
          T u = vx[j]; vx[j] = vy[j]; vy[j] = u;
          if constexpr (CPLX) { u = vxi[j]; vxi[j] = vyi[j]; vyi[j] = u; }
        }
      }
    }
    // an update that overflowed anywhere in the group leaves G unstored
    const unsigned long long gmask = (lg == 6) ? ~0ull : (((1ull << LP) - 1ull) << (g << lg));
    int err = D.err;
    if (!err && ((__ballot(!isfinite_(tG)) & gmask) != 0ull))
      err = CPLX ? -630 : -580;
    // a preset V is tracked as djrot_ tracks it (an overflowed update fails)
    if (S.vtrack && !err && ((__ballot(!isfinite_(tV)) & gmask) != 0ull))
      err = CPLX ? -638 : -588;
    const unsigned dp = dsw ? q : p, dq = dsw ? p : q;
    const bool stG = valid && !err && ((code != 1) || S.sG);
    const bool stV = valid && (S.sV || (ok && (code != 1) && (code != 3)));
    if (stG)
      store_cols<T, CPLX, E>(Gr, Gi, m, nE, dp, dq, t, LP, xr, yr, xi, yi);
    if (stV)
      store_cols<T, CPLX, E>(Vr, Vi, n, nEV, dp, dq, t, LP, vx, vy, vxi, vyi);
    if (wxf)
      *wxf += 2ull * PL *
              ((unsigned long long)__popcll(__ballot(gmem && stG && (t == 0))) * (unsigned long long)m +
               (unsigned long long)__popcll(__ballot(stV && (t == 0))) * (unsigned long long)n);
    // (e,f) of the pair as pair_apply leaves them: exchanged for -1 / -3
    if (valid && !err && (t == 0)) {
      const bool xe = (code == -1) || (code == -3);
      eS[p] = xe ? Pd.ey : Pd.ex;
      fS[p] = xe ? Pd.fy : Pd.fx;
      eS[q] = xe ? Pd.ex : Pd.ey;
      fS[q] = xe ? Pd.fx : Pd.fy;
    }
    const unsigned long long eb = __ballot(valid && (err != 0));
    if (eb)
      werr = __builtin_amdgcn_readlane(err, (int)__builtin_ctzll(eb));
    if (valid && !err) {
      wG = xmax(wG, tG);
      if (S.vtrack)
        wV = xmax(wV, tV);
    }
    wcnt += (unsigned long long)__popcll(__ballot(valid && !err && (t == 0) && D.trans));
    if constexpr (PF) {
      p = pn;
      q = qn;
      valid = validn;
#pragma unroll
      for (int j = 0; j < E; ++j) {
        xr[j] = nxr[j]; yr[j] = nyr[j]; xi[j] = nxi[j]; yi[j] = nyi[j];
        vx[j] = nvx[j]; vy[j] = nvy[j]; vxi[j] = nvxi[j]; vyi[j] = nvyi[j];
      }
    }
  }
}

// Short columns (the short classes, m <= 128 real, m <= 64 complex): 2^LG
// lanes per pair (16, 8 or 4), 64 / 2^LG pairs per wave, with the chunks
// per lane that 64 rows take, or twice as many beyond 64 rows.  The narrower
// the group, the more pairs share one instruction stream of reductions and
// decisions, and the more rows each lane updates.  (Groups wider than a row
// lose to the two-pairs-per-wave form of wave_step, whose reductions scatter
// several sums over the wave.)
template <typename T, bool CPLX, int E, int W, int LG>
__device__ __forceinline__ void group_dispatch(T *Gr, T *Gi, T *Vr, T *Vi, T *eS, T *fS, const int m, const int n,
                                               const int ord, const unsigned st, const int wave, const int lane,
                                               const Step<T> &S, T &wG, T &wV, unsigned long long &wcnt, int &werr,
                                               unsigned long long &wev, unsigned long long *wxf, const bool gmem)
{
  constexpr int E64 = 64 >> LG; // chunks of 2^LG rows for 64 rows
  if ((E < 2) || CPLX || (m <= 64)) {
    group_step<T, CPLX, E64, W, GroupShape<T, CPLX, E64>::PF, LG>(Gr, Gi, Vr, Vi, eS, fS, m, n, ord, st, wave, lane, S,
                                                                  wG, wV, wcnt, werr, wev, wxf, gmem);
    return;
  }
  if constexpr ((E >= 2) && !CPLX && (2 * E64 <= 16))
    group_step<T, CPLX, 2 * E64, W, GroupShape<T, CPLX, 2 * E64>::PF, LG>(Gr, Gi, Vr, Vi, eS, fS, m, n, ord, st, wave,
                                                                          lane, S, wG, wV, wcnt, werr, wev, wxf,
                                                                          gmem);
}

// Longer columns: a pair per 64 lanes, two pairs per wave, with the register
// footprint (64-row chunks per lane) of this matrix's row count.
template <typename T, bool CPLX, int E, int W>
__device__ __forceinline__ void wave_dispatch(T *Gr, T *Gi, T *Vr, T *Vi, T *eS, T *fS, const int m, const int n,
                                              const int ord, const unsigned st, const int wave, const int lane,
                                              const Step<T> &S, T &wG, T &wV, unsigned long long &wcnt, int &werr,
                                              int *mod, const int gstep, unsigned long long &wev,
                                              unsigned long long &wxf)
{
  const int nE = (m + 63) >> 6;
#define VJ_WS(EE)                                                                                                     \
  wave_step<T, CPLX, EE, W, StepShape<T, CPLX, EE>::GP, StepShape<T, CPLX, EE>::PF>(                                   \
      Gr, Gi, Vr, Vi, eS, fS, m, n, nE, ord, st, wave, lane, S, wG, wV, wcnt, werr, mod, gstep, wev, wxf)
  if ((E < 4) || (nE <= 2)) {
    constexpr int E2 = (E >= 2) ? 2 : E;
    VJ_WS(E2);
    return;
  }
  if constexpr (E >= 4) {
    if ((E < 8) || (nE <= 4)) {
      VJ_WS(4);
      return;
    }
  }
  if constexpr (E >= 8) {
    if ((E < 16) || (nE <= 8)) {
      VJ_WS(8);
      return;
    }
  }
  if constexpr (E >= 16)
    VJ_WS(E);
#undef VJ_WS
}

// Block rounds of the wave form (real data, pjs id 6), with the register
// footprint of this matrix's row count: two units at a time up to 4 row
// chunks (block_round2), one prefetched or single unit beyond.
template <typename T, int E, int W>
__device__ __forceinline__ void block_dispatch(T *Gr, T *Vr, T *eS, T *fS, const int m, const int n, const unsigned s,
                                               const int wave, const int lane, const Step<T> &S, T &wG,
                                               unsigned long long &wcnt, int &werr, int *mod, const int g0,
                                               unsigned long long &wev, unsigned long long &wxf)
{
  const int nE = (m + 63) >> 6;
#define VJ_BR(EE)                                                                                                     \
  block_round<T, EE, W, BlockShape<T, EE>::PF>(Gr, Vr, eS, fS, m, n, nE, s, wave, lane, S, wG, wcnt, werr, mod, g0,  \
                                               wev, wxf)
#define VJ_BR2(EE)                                                                                                    \
  block_round2<T, EE, W>(Gr, Vr, eS, fS, m, n, nE, s, wave, lane, S, wG, wcnt, werr, mod, g0, wev, wxf)
  if ((E < 4) || (nE <= 2)) {
    constexpr int E2 = (E >= 2) ? 2 : E;
    VJ_BR2(E2);
    return;
  }
  if constexpr (E >= 4) {
    if ((E < 8) || (nE <= 4)) {
      VJ_BR2(4);
      return;
    }
  }
  if constexpr (E >= 8) {
    if ((E < 16) || (nE <= 8)) {
      VJ_BR(8);
      return;
    }
  }
  if constexpr (E >= 16)
    VJ_BR(E);
#undef VJ_BR
#undef VJ_BR2
}

// Minimum waves per SIMD the register allocation must allow: two, which keeps
// the double-buffered column pairs in registers, except for the widest
// complex pairs, which would spill at 256 registers and run one wave per
// SIMD instead.
template <typename T, bool CPLX, int E> struct BsvdMinWaves {
  static constexpr int value =
      (CPLX ? (E * (int)sizeof(T) * ((sizeof(T) == 8) ? 2 : 1) >= 32) : (E * (int)sizeof(T) >= 128)) ? 1 : 2;
};

// LDS-resident form (LDS = true, lane-group steps only): the setup scan
// copies the matrix's G into dynamic LDS, every step of every sweep reads
// and writes that copy, and G (U after the epilogue) is copied out once at
// the end, so a step's G latency is that of LDS rather than of a round trip
// through the caches.  V (n x n, at most a third of the matrix's elements
// for m >= 2n) stays in device memory, where it is read and written through
// the caches: the LDS a matrix takes is that of G alone, and more matrices
// are resident per CU.
// LG (lane-group form): 2^LG lanes per pair.
template <typename T, bool CPLX, int E, int W, bool GRP, bool LDS, int LG>
__global__ __launch_bounds__(64 * W, (BsvdMinWaves<T, CPLX, E>::value)) void k_bsvd(const BatchArgs<T> a)
{
  static_assert(!LDS || GRP, "the LDS-resident form uses the lane-group steps");
  extern __shared__ double s_dyn[];            // LDS form: G (re, im)
  __shared__ int s_b;
  __shared__ T s_eS[128], s_fS[128];           // the matrix's (e,f) column norms (n <= 128)
  __shared__ int s_mod[128];                   // step at which each column was last modified
  __shared__ unsigned long long s_ev;          // pair evaluations of the matrix
  __shared__ unsigned long long s_xf;          // elements of G and V moved to and from memory
  __shared__ double s_max[2][W][2];            // per-wave (tG, tV) maxima, by step parity
  __shared__ unsigned long long s_cnt[2][W];   // per-wave transformation counts, by step parity
  __shared__ int s_err[2][W];
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
  const int rot = CPLX ? Lim<T>::ROT_EXP_C - 1 : Lim<T>::ROT_EXP_R;

  for (;;) {
    __syncthreads();
    if (threadIdx.x == 0)
      s_b = (int)atomicAdd(a.queue, 1u);
    __syncthreads();
    const int bi = __builtin_amdgcn_readfirstlane(s_b);
    if (bi >= a.nb)
      break;
    const int b = a.order[bi];
    const int m = (int)a.m[b], n = (int)a.n[b];
    T *const gGr = a.Gr + a.offG[b];
    T *const gGi = CPLX ? a.Gi + a.offG[b] : nullptr;
    T *const gVr = a.Vr + a.offV[b];
    T *const gVi = CPLX ? a.Vi + a.offV[b] : nullptr;
    T *const lG = (T *)s_dyn;
    T *const Gr = LDS ? lG : gGr;
    T *const Gi = CPLX ? (LDS ? lG + m * n : gGi) : nullptr;
    T *const Vr = gVr;
    T *const Vi = CPLX ? gVi : nullptr;
    T *const eS = a.eS + a.offS[b];
    T *const fS = a.fS + a.offS[b];
    const unsigned stp = (unsigned)n; // steps per sweep (pjs id 4 or 6)
    unsigned long long wev = 0ull;    // this wave's pair evaluations
    unsigned long long wxf = 0ull;    // this wave's elements moved (global-memory forms)
    if (threadIdx.x == 0) {
      s_ev = 0ull;
      s_xf = 0ull;
    }
    // The block form runs the two steps of a round without a fold between
    // them, which equals the step-by-step run as long as no step rescales G
    // or V.  For real data none can: G starts scaled to max |g| < 2^NRM_EXP
    // <= M/(2m) (M the largest finite value), the update maxima are bounded
    // by |g_ij| <= |G|_F <= sqrt(m n) max |g| < M/2, i.e. below 2^(ROT_EXP_R
    // + 1) with a margin of at least 2^-10 relative (m <= 1024), which the
    // rounding of the rotations cannot close; V stays orthogonal.
    const bool blk = !GRP && !CPLX && (a.ord == 6) && !a.vpreset;

    // ---- setup (dvjsvd.c:78-184): max |G|, V := I, eS/fS, initial scale
    T mxv = (T)0;
    for (long long t = threadIdx.x; t < (long long)m * n; t += 64 * W) {
      const T g = gGr[t];
      mxv = xmax(xmin(vabs(g), inf_<T>()), mxv);
      if constexpr (LDS)
        Gr[t] = g;
      if constexpr (CPLX) {
        const T gi = gGi[t];
        mxv = xmax(xmin(vabs(gi), inf_<T>()), mxv);
        if constexpr (LDS)
          Gi[t] = gi;
      }
    }
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1)
      mxv = xmax(mxv, __shfl_xor(mxv, o, 64));
    if (lane == 0)
      s_max[0][wave][0] = (double)mxv;
    __syncthreads();
    T MG = (T)s_max[0][0][0];
    for (int w = 1; w < W; ++w)
      MG = xmax(MG, (T)s_max[0][w][0]);
    __syncthreads();
    // a preset V (dvjsvd.c:89-96): its max |.|
    T MV0 = (T)1;
    if (a.vpreset) {
      T mv = (T)0;
      for (int t = threadIdx.x; t < n * n; t += 64 * W) {
        const T v = gVr[t];
        mv = xmax(xmin(vabs(v), inf_<T>()), mv);
        if constexpr (CPLX) {
          const T vi = gVi[t];
          mv = xmax(xmin(vabs(vi), inf_<T>()), mv);
        }
      }
#pragma unroll
      for (int o = 32; o >= 1; o >>= 1)
        mv = xmax(mv, __shfl_xor(mv, o, 64));
      if (lane == 0)
        s_max[1][wave][0] = (double)mv;
      __syncthreads();
      MV0 = (T)s_max[1][0][0];
      for (int w = 1; w < W; ++w)
        MV0 = xmax(MV0, (T)s_max[1][w][0]);
      __syncthreads();
    }
    int status = 0;
    if (!(MG <= Lim<T>::MAX))
      status = CPLX ? -92 : -80;
    else if (a.vpreset && !(MV0 <= Lim<T>::MAX))
      status = CPLX ? -104 : -92;
    else if (a.vpreset && !(MV0 > (T)0))
      status = CPLX ? -17 : -13;
    if (!status) {
      if (!a.vpreset)
        for (int t = threadIdx.x; t < n * n; t += 64 * W) {
          const int j = t / n, i = t - j * n;
          Vr[t] = (i == j) ? (T)1 : (T)0;
          if constexpr (CPLX)
            Vi[t] = (T)0;
        }
      for (int j = threadIdx.x; j < n; j += 64 * W) {
        s_eS[j] = -inf_<T>();
        s_fS[j] = (T)1;
        s_mod[j] = 0;
      }
    }
    const bool norms = !status; // eS/fS are written back (from LDS) at the end
    // elements the global-memory forms move outside the steps: the scan of
    // G, V := I, and below the initial scaling and the epilogue
    unsigned long long xset = (unsigned long long)m * (unsigned long long)n + (norms ? (unsigned long long)n * n : 0ull);
    unsigned sw = 0;
    unsigned long long total = 0;
    int sT = 0, NRM_EXP = 0, sVt = 0; // sVt: the shifts of a preset V
    T MV = (T)1;
    if (!status && (MG > (T)0)) {
      const T M_m = CPLX ? (Lim<T>::MAX / ((T)(m << 2) * (T)1.41421356237309504880)) : (Lim<T>::MAX / (T)(m << 1));
      NRM_EXP = frexp_exp(M_m) - 1;
      const int eM = frexp_exp(MG) - 1;
      const int sN = NRM_EXP - eM - 1;
      MG = ldexp_(MG, sN);
      sT = sN;
      if (sN) {
        for (long long t = threadIdx.x; t < (long long)m * n; t += 64 * W) {
          Gr[t] = ldexp_(Gr[t], sN);
          if constexpr (CPLX)
            Gi[t] = ldexp_(Gi[t], sN);
        }
        xset += 2ull * (unsigned long long)m * (unsigned long long)n;
      }
      if (a.vpreset) {
        // its initial scaling (dvjsvd.c:143-156): max |V| to 2^rot
        const int eV = frexp_exp(MV0) - 1;
        sVt = rot - eV;
        MV = ldexp_(MV0, sVt);
        if (sVt) {
          for (int t = threadIdx.x; t < n * n; t += 64 * W) {
            Vr[t] = ldexp_(Vr[t], sVt);
            if constexpr (CPLX)
              Vi[t] = ldexp_(Vi[t], sVt);
          }
          xset += 2ull * (unsigned long long)n * (unsigned long long)n;
        }
      }
      __syncthreads();
      Step<T> S;
      S.tol = sqrt_((T)m) * (Lim<T>::EPS / (T)2);
      S.gst = ldexp_(S.tol, (int)Lim<T>::FIN_EXP);
      S.sG = 0;
      S.sV = 0;
      S.vtrack = a.vpreset ? 1 : 0;
      S.sec = 0;
      int par = 0, gstep = 0;
      for (sw = 0; sw < a.swp;) {
        unsigned long long swt = 0;
        for (unsigned st = 0; st < stp; ++st) {
          ++gstep;
          T wG = (T)0, wV = (T)0;
          unsigned long long wcnt = 0;
          int werr = 0;
          if constexpr (GRP)
            group_dispatch<T, CPLX, E, W, LG>(Gr, Gi, Vr, Vi, s_eS, s_fS, m, n, a.ord, st, wave, lane, S, wG, wV,
                                              wcnt, werr, wev, &wxf, !LDS);
          else {
            if constexpr (!CPLX) {
              if (blk) {
                block_dispatch<T, E, W>(Gr, Vr, s_eS, s_fS, m, n, st >> 1, wave, lane, S, wG, wcnt, werr, s_mod,
                                        gstep, wev, wxf);
                ++gstep;
                ++st;
              }
            }
            if (!blk)
              wave_dispatch<T, CPLX, E, W>(Gr, Gi, Vr, Vi, s_eS, s_fS, m, n, a.ord, st, wave, lane, S, wG, wV, wcnt,
                                           werr, s_mod, gstep, wev, wxf);
          }
          wG = red_max<T, 64>(wG, nullptr);
          if (S.vtrack)
            wV = red_max<T, 64>(wV, nullptr);
          if (lane == 0) {
            s_max[par][wave][0] = (double)wG;
            s_max[par][wave][1] = (double)wV;
            s_cnt[par][wave] = wcnt;
            s_err[par][wave] = werr;
          }
          __syncthreads();
          // every thread folds the same LDS values: uniform decisions
          T nMG = (T)0, nMV = (T)0;
          unsigned long long cnt = 0;
          int err = 0;
          for (int w = 0; w < W; ++w) {
            nMG = xmax(nMG, (T)s_max[par][w][0]);
            nMV = xmax(nMV, (T)s_max[par][w][1]);
            cnt += s_cnt[par][w];
            err = err ? err : s_err[par][w];
          }
          par ^= 1;
          if (err) {
            status = err;
            break;
          }
          swt += cnt;
          MG = (nMG > MG) ? nMG : MG;
          MV = (nMV > MV) ? nMV : MV;
          S.sG = 0;
          S.sV = 0;
          {
            const int eM = frexp_exp(MG) - 1;
            if (rot - eM < 0) {
              S.sG = NRM_EXP - eM - 1;
              MG = ldexp_(MG, S.sG);
              sT += S.sG;
            }
            const int eV = frexp_exp(MV) - 1;
            if (rot - eV < 0) {
              S.sV = rot - eV;
              MV = ldexp_(MV, S.sV);
              sVt += S.sV;
            }
          }
          __threadfence_block();
        }
        if (status)
          break;
        total += swt;
        if (!swt)
          break;
        ++sw;
      }
      __syncthreads();
      // ---- epilogue (dvjsvd.c:604-631); V is rescaled only when preset (V
      // := I is never rescaled: it stays orthogonal)
      if (!status && (sw < a.swp)) {
        for (long long t = threadIdx.x; t < (long long)m * n; t += 64 * W) {
          const int j = (int)(t / m);
          const T f_ = s_fS[j], s_ = -s_eS[j];
          Gr[t] = scalef(Gr[t] / f_, s_);
          if constexpr (CPLX)
            Gi[t] = scalef(Gi[t] / f_, s_);
        }
        if (sVt) {
          const T sv = (T)(-sVt);
          for (int t = threadIdx.x; t < n * n; t += 64 * W) {
            Vr[t] = scalef(Vr[t], sv);
            if constexpr (CPLX)
              Vi[t] = scalef(Vi[t], sv);
          }
          xset += 2ull * (unsigned long long)n * (unsigned long long)n;
        }
        __syncthreads();
        for (int j = threadIdx.x; j < n; j += 64 * W)
          s_eS[j] -= (T)sT;
        xset += 2ull * (unsigned long long)m * (unsigned long long)n;
      }
    }
    if constexpr (LDS) {
      // G (U after the epilogue, else the working G) out of LDS
      __syncthreads();
      if (norms) {
        for (long long t = threadIdx.x; t < (long long)m * n; t += 64 * W) {
          gGr[t] = Gr[t];
          if constexpr (CPLX)
            gGi[t] = Gi[t];
        }
      }
    }
    if (lane == 0 && wev)
      atomicAdd(&s_ev, wev);
    if (lane == 0 && wxf)
      atomicAdd(&s_xf, wxf);
    __syncthreads();
    if (norms)
      for (int j = threadIdx.x; j < n; j += 64 * W) {
        eS[j] = s_eS[j];
        fS[j] = s_fS[j];
      }
    if (threadIdx.x == 0) {
      a.info[b] = status ? (long long)status : (long long)sw;
      if (a.transforms)
        a.transforms[b] = total;
      if (a.evals)
        a.evals[b] = s_ev;
      if (a.xfer) {
        // the LDS-resident form moves G in and U (or G) out once, and V
        // (formed, and read and written by the steps) through device memory
        const unsigned long long pl = CPLX ? 2ull : 1ull;
        a.xfer[b] = LDS ? s_xf + pl * (unsigned long long)n * (unsigned long long)(norms ? (2 * m + n) : m)
                        : (s_xf + pl * xset);
      }
    }
  }
}

// Dynamic LDS of a launch: the resident matrices of the LDS form.
template <typename T, bool CPLX, int E, int W, bool GRP, bool LDS>
static size_t launch_lds(const BatchArgs<T> &a)
{
  return LDS ? (size_t)a.lds_bytes : 0;
}

template <typename T, bool CPLX, int E, int W, bool GRP, bool LDS, int LG>
static int launch_bsvd(const BatchArgs<T> &a, int grid, hipStream_t s)
{
  const size_t lds = launch_lds<T, CPLX, E, W, GRP, LDS>(a);
  if (lds > 64 * 1024)
    VJ_HIP(hipFuncSetAttribute((const void *)k_bsvd<T, CPLX, E, W, GRP, LDS, LG>,
                               hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
  hipLaunchKernelGGL((k_bsvd<T, CPLX, E, W, GRP, LDS, LG>), dim3(grid), dim3(64 * W), lds, s, a);
  return check_launch();
}

template <typename T, bool CPLX, int E, int W, bool GRP, bool LDS, int LG>
static int occupancy(int *blocks, size_t lds)
{
  const hipError_t e =
      hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks, k_bsvd<T, CPLX, E, W, GRP, LDS, LG>, 64 * W, lds);
  return (e == hipSuccess) ? 0 : set_error(VJ_EHIP - (int)e, "occupancy query failed");
}

template <typename T, bool CPLX, int W, bool GRP, bool LDS, int LG>
static int dispatch_E(const BatchArgs<T> &a, int max_m, int grid_cap, hipStream_t s)
{
  int cus = 0, blocks = 1, dev = 0;
  VJ_HIP(hipGetDevice(&dev));
  VJ_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
#define VJ_BS_CASE(EE)                                                         \
  {                                                                            \
    const size_t lds = launch_lds<T, CPLX, EE, W, GRP, LDS>(a);                \
    if (int rc = occupancy<T, CPLX, EE, W, GRP, LDS, LG>(&blocks, lds))        \
      return rc;                                                               \
    long long g = (long long)cus * (blocks > 0 ? blocks : 1);                  \
    if (grid_cap > 0 && g > grid_cap) g = grid_cap;                            \
    if (g > a.nb) g = a.nb;                                                    \
    return launch_bsvd<T, CPLX, EE, W, GRP, LDS, LG>(a, (int)g, s);            \
  }
  if (max_m <= 64) VJ_BS_CASE(1)
  if constexpr (!GRP || group_shape_ok(CPLX, 128, LG, W))
    if (max_m <= 128) VJ_BS_CASE(2)
  if constexpr (!GRP) {
    if (max_m <= 256) VJ_BS_CASE(4)
    if (max_m <= 512) VJ_BS_CASE(8)
    if (max_m <= 1024) VJ_BS_CASE(16)
  }
#undef VJ_BS_CASE
  return set_error(-1, "batched SVD: m = %d exceeds %d", max_m, GRP ? (int)group_max_rows(CPLX, LG) : 1024);
}

template <typename T, bool CPLX, int W, bool LDS, int LG>
static int dispatch_LG(const BatchArgs<T> &a, int max_m, int grid_cap, hipStream_t s)
{
  if constexpr (group_shape_ok(CPLX, 64, LG, W))
    return dispatch_E<T, CPLX, W, true, LDS, LG>(a, max_m, grid_cap, s);
  return set_error(-1, "batched SVD: no lane-group form with %d lanes per pair and %d waves", 1 << LG, W);
}

template <typename T, bool CPLX, int W, bool LDS>
static int dispatch_grp(const BatchArgs<T> &a, int max_m, int grid_cap, hipStream_t s)
{
  if (max_m > group_max_rows(CPLX, a.group_lg))
    return set_error(-1, "batched SVD: m = %d exceeds the %d rows of the lane-group form with %d lanes per pair",
                     max_m, group_max_rows(CPLX, a.group_lg), 1 << a.group_lg);
  switch (a.group_lg) {
  case 4: return dispatch_LG<T, CPLX, W, LDS, 4>(a, max_m, grid_cap, s);
  case 3: return dispatch_LG<T, CPLX, W, LDS, 3>(a, max_m, grid_cap, s);
  case 2: return dispatch_LG<T, CPLX, W, LDS, 2>(a, max_m, grid_cap, s);
  default: return set_error(-1, "batched SVD: %d lanes per pair (16, 8 or 4)", 1 << a.group_lg);
  }
}

template <typename T, bool CPLX, int W>
static int dispatch_G(const BatchArgs<T> &a, int max_m, int grid_cap, bool grp, hipStream_t s)
{
  if (!grp)
    return dispatch_E<T, CPLX, W, false, false, 4>(a, max_m, grid_cap, s);
  if (a.lds_bytes > 0) {
    if (a.lds_bytes > LDS_FORM_MAX_BYTES)
      return set_error(-1, "batched SVD: %d bytes of LDS per matrix exceed %d", a.lds_bytes, LDS_FORM_MAX_BYTES);
    return dispatch_grp<T, CPLX, W, true>(a, max_m, grid_cap, s);
  }
  return dispatch_grp<T, CPLX, W, false>(a, max_m, grid_cap, s);
}

template <typename T, bool CPLX>
int batched_svd(const BatchArgs<T> &a, int waves, int max_m, int grid_cap, bool grp, hipStream_t s)
{
  if (a.nb <= 0)
    return 0;
  switch (waves) {
  case 1: return dispatch_G<T, CPLX, 1>(a, max_m, grid_cap, grp, s);
  case 2: return dispatch_G<T, CPLX, 2>(a, max_m, grid_cap, grp, s);
  case 4: return dispatch_G<T, CPLX, 4>(a, max_m, grid_cap, grp, s);
  case 8: return dispatch_G<T, CPLX, 8>(a, max_m, grid_cap, grp, s);
  default: return set_error(-1, "batched SVD: waves per workgroup must be 1, 2, 4 or 8 (got %d)", waves);
  }
}

} // namespace vj
