// One Jacobi step on one column pair, in three phases, executed by a group of
// B lanes (one wave for B = 64, a whole workgroup for B = 256/1024):
//
//   pair_reduce  exact column norms in (e,f) form and the scaled cosine,
//                from the pair held in registers (E elements per lane, row
//                lane + j*B):
//                  - max-|.| exponent scaling, double-double sums of squares
//                    and a corrected square root (replaces dnorm2_/znorm2_,
//                    dnorm2.c:9-14, znorm2.c:9-15);
//                  - the scaled, compensated dot product (replaces
//                    ddpscl_/zdpscl_, ddpscl.c:7-75, zdpscl.c:21-116):
//                    TwoProd + TwoSum per element and a double-double
//                    reduce-scatter across the lanes.
//   pair_decide  per-pair scalar work, one lane per pair: the convergence
//                test, Gram prescaling and action code (dvjsvd.c:349-459)
//                and the 2x2 EVD (dbjac2i cosine form).
//   pair_apply   the in-place update: rotation (djrot_/djrotf_/zjrot_/
//                zjrotf_), swap (dswp_) or Gram-Schmidt correction
//                (dgsscl_/zgsscl_) with the reference's fused multiply-add
//                sequences, and the max-|.| tracking of the overflow control.
//
// pair_step runs the three phases back to back (used where one group owns
// one pair); the batched kernel runs pair_decide for all pairs of a step in
// parallel, one lane per pair, between an LDS-staged reduce and apply.
#pragma once

#include "vj_dev.hpp"

namespace vj {

// -------------------------------------------------------------- reductions
// Wave-level exchanges without an LDS round trip: DPP moves inside rows of
// 16 lanes (mirror, half mirror, quad xor 2, quad xor 1: after the four
// levels every lane of a row holds the row total) and the gfx950 permlane
// swaps between the wave halves (32) and neighbouring rows (16):
//   swap32(a, b): a := [a lanes 0-31, b lanes 0-31], b := [a lanes 32-63, b lanes 32-63]
//   swap16(a, b): a := rows (a0, b0, a2, b2),       b := rows (a1, b1, a3, b3)
// so a + b after a swap is a reduce-scatter step of two quantities.
constexpr int DPP_MIRROR = 0x140, DPP_HMIRROR = 0x141, DPP_XOR2 = 0x4E, DPP_XOR1 = 0xB1;

template <int CTRL> __device__ __forceinline__ double dpp_(double v)
{
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_update_dpp(0, (int)b, CTRL, 0xF, 0xF, false);
  const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), CTRL, 0xF, 0xF, false);
  return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned)lo);
}
template <int CTRL> __device__ __forceinline__ float dpp_(float v)
{
  return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, 0xF, 0xF, false));
}

__device__ __forceinline__ void swap32_(unsigned &a, unsigned &b)
{
  const auto r = __builtin_amdgcn_permlane32_swap(a, b, false, false);
  a = r[0];
  b = r[1];
}
__device__ __forceinline__ void swap16_(unsigned &a, unsigned &b)
{
  const auto r = __builtin_amdgcn_permlane16_swap(a, b, false, false);
  a = r[0];
  b = r[1];
}
template <bool H32> __device__ __forceinline__ void swap_(double &a, double &b)
{
  const unsigned long long ua = (unsigned long long)__double_as_longlong(a);
  const unsigned long long ub = (unsigned long long)__double_as_longlong(b);
  unsigned al = (unsigned)ua, ah = (unsigned)(ua >> 32), bl = (unsigned)ub, bh = (unsigned)(ub >> 32);
  if constexpr (H32) {
    swap32_(al, bl);
    swap32_(ah, bh);
  }
  else {
    swap16_(al, bl);
    swap16_(ah, bh);
  }
  a = __longlong_as_double((long long)(((unsigned long long)ah << 32) | al));
  b = __longlong_as_double((long long)(((unsigned long long)bh << 32) | bl));
}
template <bool H32> __device__ __forceinline__ void swap_(float &a, float &b)
{
  unsigned ua = __float_as_uint(a), ub = __float_as_uint(b);
  if constexpr (H32)
    swap32_(ua, ub);
  else
    swap16_(ua, ub);
  a = __uint_as_float(ua);
  b = __uint_as_float(ub);
}

__device__ __forceinline__ double readlane_(double v, int l)
{
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readlane((int)b, l);
  const int hi = __builtin_amdgcn_readlane((int)(b >> 32), l);
  return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned)lo);
}
__device__ __forceinline__ float readlane_(float v, int l) { return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), l)); }

template <typename T> struct DD {
  T s, c;
};

template <typename T> __device__ __forceinline__ DD<T> dd_add(DD<T> a, DD<T> b)
{
  DD<T> r;
  T t;
  two_sum(a.s, b.s, r.s, t);
  r.c = (a.c + b.c) + t;
  return r;
}

template <typename T> __device__ __forceinline__ DD<T> dd_sel(bool c, DD<T> a, DD<T> b)
{
  return DD<T>{c ? a.s : b.s, c ? a.c : b.c};
}

// s := s + p exactly kept in (s, c)
template <typename T> __device__ __forceinline__ void dd_acc(DD<T> &a, T p, T e)
{
  T t;
  two_sum(a.s, p, a.s, t);
  a.c += (t + e);
}

template <int CTRL, typename T> __device__ __forceinline__ DD<T> dd_dpp(DD<T> v)
{
  return DD<T>{dpp_<CTRL>(v.s), dpp_<CTRL>(v.c)};
}
template <bool H32, typename T> __device__ __forceinline__ DD<T> dd_swapadd(DD<T> a, DD<T> b)
{
  swap_<H32>(a.s, b.s);
  swap_<H32>(a.c, b.c);
  return dd_add(a, b);
}

// totals (every lane of an aligned group of LP = 16, 8 or 4 lanes of a row)
// of a max / a double-double sum: the mirror pairs the halves of a row, the
// half mirror the halves of an 8-lane group, the quad exchanges the rest
template <typename T, int LP = 16> __device__ __forceinline__ T row_max(T v)
{
  static_assert(LP == 16 || LP == 8 || LP == 4, "row groups of 16, 8 or 4 lanes");
  if constexpr (LP >= 16)
    v = xmax(v, dpp_<DPP_MIRROR>(v));
  if constexpr (LP >= 8)
    v = xmax(v, dpp_<DPP_HMIRROR>(v));
  v = xmax(v, dpp_<DPP_XOR2>(v));
  v = xmax(v, dpp_<DPP_XOR1>(v));
  return v;
}
template <typename T, int LP = 16> __device__ __forceinline__ DD<T> row_sum(DD<T> v)
{
  static_assert(LP == 16 || LP == 8 || LP == 4, "row groups of 16, 8 or 4 lanes");
  if constexpr (LP >= 16)
    v = dd_add(v, dd_dpp<DPP_MIRROR>(v));
  if constexpr (LP >= 8)
    v = dd_add(v, dd_dpp<DPP_HMIRROR>(v));
  v = dd_add(v, dd_dpp<DPP_XOR2>(v));
  v = dd_add(v, dd_dpp<DPP_XOR1>(v));
  return v;
}

// Broadcast of row values over the wave, kept in VGPRs (readlane would
// make them scalar and crowd the SGPR file): k holds k_r in every lane of
// row r; afterwards q_r holds k_r in every lane.
template <typename T> __device__ __forceinline__ void bcast_rows4(const T k, T &q0, T &q1, T &q2, T &q3)
{
  T a = k, b = k;
  swap_<true>(a, b);     // a: rows (k0, k1, k0, k1), b: rows (k2, k3, k2, k3)
  q0 = a;
  q1 = a;
  swap_<false>(q0, q1);  // q0: k0 everywhere, q1: k1 everywhere
  q2 = b;
  q3 = b;
  swap_<false>(q2, q3);
}
template <typename T> __device__ __forceinline__ void bcast_rows4(const DD<T> k, DD<T> &q0, DD<T> &q1, DD<T> &q2, DD<T> &q3)
{
  bcast_rows4(k.s, q0.s, q1.s, q2.s, q3.s);
  bcast_rows4(k.c, q0.c, q1.c, q2.c, q3.c);
}

// Max of two values over the group.  Wave level: one swap32 scatters a to
// the low half and b to the high half, the rows reduce with DPP, rows 0/1
// and 2/3 combine through one swap16 and a last swap32 spreads a and b.
// The group reductions below write lds and read it back after a barrier;
// PRE = false leaves out the barrier before the write, for an lds area
// that no lane can still be reading.
template <typename T, int B, bool PRE = true>
__device__ __forceinline__ void red_max2(T &a, T &b, double *lds)
{
  T lo = a, hi = b;
  swap_<true>(lo, hi);
  const T k = row_max(xmax(lo, hi));
  T c = k, d = k;
  swap_<false>(c, d);    // c: rows (k0, k0, k2, k2), d: rows (k1, k1, k3, k3)
  T x = xmax(c, d), y = x;
  swap_<true>(x, y);     // x: a everywhere, y: b everywhere
  a = x;
  b = y;
  if constexpr (B > 64) {
    constexpr int W = B / 64;
    const int lane = threadIdx.x & 63;
    const int w = threadIdx.x >> 6;
    if constexpr (PRE)
      __syncthreads();
    if (lane == 0) {
      lds[2 * w] = (double)a;
      lds[2 * w + 1] = (double)b;
    }
    __syncthreads();
    T ra = (T)lds[0], rb = (T)lds[1];
#pragma unroll
    for (int k2 = 1; k2 < W; ++k2) {
      ra = xmax(ra, (T)lds[2 * k2]);
      rb = xmax(rb, (T)lds[2 * k2 + 1]);
    }
    a = ra;
    b = rb;
  }
}

// Sum of two double-double values over a wave (64 lanes).  Each value goes
// through the same additions, in the same order, as a value of red_dd4: lane
// l plus lane l + 32 (swap32), the two rows of each half (swap16), then the
// row reduction; so its totals are bit-identical to red_dd4's, at half the
// exchanges and additions.
template <typename A> __device__ __forceinline__ void red_dd2(DD<A> &q0, DD<A> &q1)
{
  const DD<A> k = dd_swapadd<true>(q0, q1);          // low half: q0 partials, high half: q1 partials
  const DD<A> r = row_sum(dd_swapadd<false>(k, k));  // rows 0, 1: q0; rows 2, 3: q1
  DD<A> a = r, b = r;
  swap_<true>(a.s, b.s);                             // a: rows (0, 1, 0, 1), b: rows (2, 3, 2, 3)
  swap_<true>(a.c, b.c);
  q0 = a;
  q1 = b;
}

// Sum of four double-double values over the group.  Wave level: swap32
// scatters (q0, q2) and (q1, q3) over the halves, swap16 scatters the two
// results over the rows (row r then holds a partial sum of q_r), the rows
// reduce with DPP and the row totals are spread to every lane.
template <typename A, int B, bool PRE = true>
__device__ __forceinline__ void red_dd4(DD<A> &q0, DD<A> &q1, DD<A> &q2, DD<A> &q3, double *lds)
{
  const DD<A> k0 = dd_swapadd<true>(q0, q2);
  const DD<A> k1 = dd_swapadd<true>(q1, q3);
  const DD<A> k = row_sum(dd_swapadd<false>(k0, k1));
  bcast_rows4(k, q0, q1, q2, q3);
  if constexpr (B > 64) {
    constexpr int W = B / 64;
    const int lane = threadIdx.x & 63;
    const int w = threadIdx.x >> 6;
    if constexpr (PRE)
      __syncthreads();
    if (lane == 0) {
      double *o = lds + 8 * w;
      o[0] = (double)q0.s; o[1] = (double)q0.c;
      o[2] = (double)q1.s; o[3] = (double)q1.c;
      o[4] = (double)q2.s; o[5] = (double)q2.c;
      o[6] = (double)q3.s; o[7] = (double)q3.c;
    }
    __syncthreads();
    DD<A> r0{(A)lds[0], (A)lds[1]}, r1{(A)lds[2], (A)lds[3]}, r2{(A)lds[4], (A)lds[5]}, r3{(A)lds[6], (A)lds[7]};
#pragma unroll
    for (int w2 = 1; w2 < W; ++w2) {
      const double *o = lds + 8 * w2;
      r0 = dd_add(r0, DD<A>{(A)o[0], (A)o[1]});
      r1 = dd_add(r1, DD<A>{(A)o[2], (A)o[3]});
      r2 = dd_add(r2, DD<A>{(A)o[4], (A)o[5]});
      r3 = dd_add(r3, DD<A>{(A)o[6], (A)o[7]});
    }
    q0 = r0; q1 = r1; q2 = r2; q3 = r3;
  }
}

// max over the group (used for the update maxima)
template <typename T, int B, bool PRE = true>
__device__ __forceinline__ T red_max(T v, double *lds)
{
  const T k = row_max(v);
  T c = k, d = k;
  swap_<false>(c, d);
  T x = xmax(c, d), y = x;
  swap_<true>(x, y);
  v = xmax(x, y);
  if constexpr (B > 64) {
    constexpr int W = B / 64;
    const int w = threadIdx.x >> 6;
    if constexpr (PRE)
      __syncthreads();
    if ((threadIdx.x & 63) == 0)
      lds[w] = (double)v;
    __syncthreads();
    T r = (T)lds[0];
#pragma unroll
    for (int k2 = 1; k2 < W; ++k2)
      r = xmax(r, (T)lds[k2]);
    v = r;
  }
  return v;
}

// Running maximum of |updated element| for updates with one product per
// element (real rotations and Gram-Schmidt steps, complex rotations with a
// real tangent): finite columns and finite parameters give finite values or
// +-inf, never NaN, and there the IEEE maximum (one instruction) equals the
// reference's ordered maximum.  Updates that sum two products can make
// inf - inf and keep xmax.
template <typename T> __device__ __forceinline__ T max1p(T acc, T v) { return __builtin_fmax(acc, v); }

// ------------------------------------------------------------- (e,f) norms
// Accumulator type of the norm and dot reductions: double-double for FP64
// data, plain double for FP32 data (FP32 products are exact in double).
template <typename T> struct Acc;
template <> struct Acc<double> { typedef double A; static constexpr bool DDACC = true; };
template <> struct Acc<float> { typedef double A; static constexpr bool DDACC = false; };

// sqrt of a double-double, rounded once (one Newton correction on the
// residual computed with an FMA)
__device__ __forceinline__ double dd_sqrt(DD<double> v)
{
  double hi, lo;
  two_sum(v.s, v.c, hi, lo);
  const double r = sqrt_(hi);
  if (!(r > 0.0))
    return r;
  const double d = fma_(-r, r, hi) + lo;
  return r + d / (2.0 * r);
}

// (e,f) of ||x|| = r * 2^k with r the scaled norm (r > 0, finite)
template <typename T> __device__ __forceinline__ void ef_from(double r, int k, T &e, T &f)
{
  if (!(r > 0.0)) {
    e = -inf_<T>();
    f = (T)1;
    return;
  }
  const T rt = (T)r;
  const int q = frexp_exp(rt) - 1;
  e = (T)(k + q);
  f = ldexp_(rt, -q);
}

// ----------------------------------------------------------- pair context
template <typename T> struct Step {
  T tol, gst;        // sqrt(m) eps/2 and tol * 2^FIN_EXP (dvjsvd.c:183-184)
  int sG, sV;        // pending power-of-two rescale of G and V in this step
  int vtrack;        // 1: V preset (djrot_ tracking), 0: V = I (djrotf_)
  int sec;           // 1: the USE_SECANTS form (secants, rotations divide)
};

// Results of pair_reduce: norms of the pair (p = x, q = y) and the cosine
// a21 = y^H x / (|x| |y|) in the reference's (q, p) order.
template <typename T> struct PairScal {
  T ex, fx, ey, fy;
  T a21r, a21i;
};

// Results of pair_decide.
template <typename T> struct PairDec {
  int code;          // action code (dvjsvd.c:438-453): 3, -3, 2, -2, -1, 1
  int trans;         // 1 if transformed (counted in the sweep total)
  int err;           // 0, or the reference's negative line code
  T c, at, sat;      // cosine (secant when sec) and (complex) tangent of the rotation
  T cosv;            // |a21|: |cos| of the angle between the two columns
  int sec;           // c is the secant: the rotations divide by it
};

// The scaling of a rotated element: the product by the cosine, or in the
// reference's USE_SECANTS form (VDMULDIV, vecdef.h:97-101) the quotient by
// the secant.
template <typename T> __device__ __forceinline__ T cscale(const T v, const T c, const int sec)
{
  return sec ? (v / c) : (v * c);
}

template <typename T, bool CPLX> struct PairOut {
  int code;
  int trans;
  int err;
  T tG, tV;          // max |updated G / V element| of this pair
};

// Per-lane pieces of pair_reduce, so that a wave can reduce several pairs
// together.  pair_maxima: partial max |x|, max |y| over this lane's rows (G
// is finite whenever a pair is reduced, checked at set-up and after every
// update, so the IEEE maximum equals the ordered maximum here).
template <typename T, bool CPLX, int E>
__device__ __forceinline__ void pair_maxima(const T (&xr)[E], const T (&yr)[E], const T (&xi)[E], const T (&yi)[E],
                                            const int nE, T &mx, T &my)
{
  mx = (T)0;
  my = (T)0;
#pragma unroll
  for (int j = 0; j < E; ++j) {
    if (j < nE) {
      mx = __builtin_fmax(mx, vabs(xr[j]));
      my = __builtin_fmax(my, vabs(yr[j]));
      if constexpr (CPLX) {
        mx = __builtin_fmax(mx, vabs(xi[j]));
        my = __builtin_fmax(my, vabs(yi[j]));
      }
    }
  }
}

// Scaling exponent of a column from its max |.| (0 for a zero column).
template <typename T> __device__ __forceinline__ int max_exp(const T mx) { return (mx > (T)0) ? (frexp_exp(mx) - 1) : 0; }

// pair_sums: adds to the lane's double-double sums of |x 2^-kx|^2,
// |y 2^-ky|^2 and x^H y 2^-(kx+ky) (real part dr, imaginary part di) over
// its rows, in row order (the caller starts them at zero).
template <typename T, bool CPLX, int E>
__device__ __forceinline__ void pair_sums(const T (&xr)[E], const T (&yr)[E], const T (&xi)[E], const T (&yi)[E],
                                          const int nE, const int kx, const int ky, DD<typename Acc<T>::A> &sx,
                                          DD<typename Acc<T>::A> &sy, DD<typename Acc<T>::A> &dr,
                                          DD<typename Acc<T>::A> &di)
{
  typedef typename Acc<T>::A A;
#pragma unroll
  for (int j = 0; j < E; ++j) {
    if (j >= nE)
      break;
    const A a = ldexp_((A)xr[j], -kx), b = ldexp_((A)yr[j], -ky);
    if constexpr (Acc<T>::DDACC) {
      A p = a * a;
      dd_acc(sx, p, fma_(a, a, -p));
      p = b * b;
      dd_acc(sy, p, fma_(b, b, -p));
      p = a * b;
      dd_acc(dr, p, fma_(a, b, -p));
    }
    else {
      sx.s += a * a;
      sy.s += b * b;
      dr.s += a * b;
    }
    if constexpr (CPLX) {
      const A c = ldexp_((A)xi[j], -kx), d = ldexp_((A)yi[j], -ky);
      if constexpr (Acc<T>::DDACC) {
        A p = c * c;
        dd_acc(sx, p, fma_(c, c, -p));
        p = d * d;
        dd_acc(sy, p, fma_(d, d, -p));
        // x^H y = (a - ic)(b + id) = (ab + cd) + i(ad - cb)
        p = c * d;
        dd_acc(dr, p, fma_(c, d, -p));
        p = a * d;
        dd_acc(di, p, fma_(a, d, -p));
        p = -(c * b);
        dd_acc(di, p, -fma_(c, b, -(c * b)));
      }
      else {
        sx.s += c * c;
        sy.s += d * d;
        dr.s += c * d;
        di.s += a * d - c * b;
      }
    }
  }
}

// pair_scal: the (e,f) norms and the cosine from the group totals.
template <typename T, bool CPLX>
__device__ __forceinline__ PairScal<T> pair_scal(const T mx, const T my, const int kx, const int ky,
                                                 const DD<typename Acc<T>::A> sx, const DD<typename Acc<T>::A> sy,
                                                 const DD<typename Acc<T>::A> dr, const DD<typename Acc<T>::A> di)
{
  typedef typename Acc<T>::A A;
  const double rx = (mx > (T)0) ? dd_sqrt(DD<double>{(double)sx.s, (double)sx.c}) : 0.0;
  const double ry = (my > (T)0) ? dd_sqrt(DD<double>{(double)sy.s, (double)sy.c}) : 0.0;
  PairScal<T> o;
  ef_from<T>(rx, kx, o.ex, o.fx);
  ef_from<T>(ry, ky, o.ey, o.fy);
  // cosine of the reference's (q, p) order (dvjsvd.c:315-319): real part
  // symmetric, imaginary part of y^H x = -(imag of x^H y); the sum is
  // rounded once and moved by the exact power of two that converts the
  // max-exponent scaling into the reference's norm-exponent scaling
  o.a21r = (T)0;
  o.a21i = (T)0;
  if ((mx > (T)0) && (my > (T)0)) {
    const int shift = (kx + ky) - (int)(o.ex + o.ey);
    const T de = o.fx * o.fy;
    o.a21r = ldexp_((T)((A)dr.s + (A)dr.c), shift) / de;
    if constexpr (CPLX)
      o.a21i = ldexp_((T)(-((A)di.s + (A)di.c)), shift) / de;
  }
  return o;
}

template <typename T, bool CPLX, int B, int E>
__device__ __forceinline__ PairScal<T>
pair_reduce(const T (&xr)[E], const T (&yr)[E], const T (&xi)[E], const T (&yi)[E], double *lds, const int nE = E)
{
  typedef typename Acc<T>::A A;
  T mx, my;
  pair_maxima<T, CPLX, E>(xr, yr, xi, yi, nE, mx, my);
  red_max2<T, B>(mx, my, lds);
  const int kx = max_exp(mx), ky = max_exp(my);
  DD<A> sx{0, 0}, sy{0, 0}, dr{0, 0}, di{0, 0};
  pair_sums<T, CPLX, E>(xr, yr, xi, yi, nE, kx, ky, sx, sy, dr, di);
  red_dd4<A, B>(sx, sy, dr, di, lds);
  return pair_scal<T, CPLX>(mx, my, kx, ky, sx, sy, dr, di);
}

// Convergence test, Gram prescaling, 2x2 EVD and action code of one pair
// (dvjsvd.c:349-459, zvjsvd.c:377-491); per lane.  sec: the 2x2 EVD returns
// the secant in c (dbjac2i with n < 0, dvjsvd.c:411-415).
template <typename T, bool CPLX>
__device__ __forceinline__ PairDec<T> pair_decide(const PairScal<T> &P, const T tol, const T gst, const int sec = 0)
{
  PairDec<T> d{1, 0, 0, (T)1, (T)0, (T)0, (T)0, sec};
  if (!isfinite_(P.a21r) || !isfinite_(P.a21i)) {
    d.err = CPLX ? -345 : -320;
    return d;
  }
  T aabs;
  if constexpr (CPLX) {
    const T x_ = vabs(P.a21r), y_ = vabs(P.a21i);
    const T m_ = xmin(x_, y_), M_ = xmax(x_, y_);
    T z = xmax(m_ / M_, (T)0);
    z = sqrt_(fma_(z, z, (T)1));
    aabs = z * M_;
  }
  else
    aabs = vabs(P.a21r);
  d.cosv = aabs;
  if (tol <= aabs) {
    d.trans = 1;
    const T e1 = P.ex, e2 = P.ey, f1 = P.fx, f2 = P.fy;
    const bool ngs = eflt(e1, e2, f1, f2);
    T maf = ngs ? f1 : f2;
    T mae = ngs ? e1 : e2;
    const T Maf = ngs ? f2 : f1;
    const T Mae = ngs ? e2 : e1;
    maf = gst * maf;
    mae = mae + getexp(maf);
    maf = getmant(maf);
    const bool cgs = eflt(mae, Mae, maf, Maf);
    // Both columns are nonzero here (a zero column has a21 = 0), so e1, e2
    // are finite and f1, f2 lie in [1, 2): the quotients lie in (1/2, 2), and
    // getexp / getmant / scalef reduce to their finite, nonzero cases.
    T f12 = f1 / f2, e12 = e1 - e2, f21 = f2 / f1, e21 = e2 - e1;
    e12 = e12 + (T)(frexp_exp(f12) - 1);
    f12 = frexp_mant(f12) * (T)2;
    e21 = e21 + (T)(frexp_exp(f21) - 1);
    f21 = frexp_mant(f21) * (T)2;
    const bool c12 = efle(e12, e21, f12, f21);
    const T Ee = c12 ? e21 : e12;
    const T dd = xmin((T)Lim<T>::FIN_EXP - Ee, (T)0);
    e12 = e12 + dd;
    e21 = e21 + dd;
    const T w0 = scalef_fin(f12, e12), w1 = scalef_fin(f21, e21);
    T l1, l2;
    bool Pm;
    if constexpr (CPLX) {
      if (sec)
        cjac2<T, JAC2_SEC, true>(w0, w1, scalef_fin(P.a21r, dd), scalef_fin(P.a21i, dd), d.c, d.at, d.sat, l1, l2, Pm);
      else
        cjac2<T, JAC2_TAN, true>(w0, w1, scalef_fin(P.a21r, dd), scalef_fin(P.a21i, dd), d.c, d.at, d.sat, l1, l2, Pm);
    }
    else {
      if (sec)
        rjac2<T, JAC2_SEC, true>(w0, w1, scalef_fin(P.a21r, dd), d.c, d.at, l1, l2, Pm);
      else
        rjac2<T, JAC2_TAN, true>(w0, w1, scalef_fin(P.a21r, dd), d.c, d.at, l1, l2, Pm);
    }
    d.code = (!ngs && cgs) ? 3 : ((ngs && cgs) ? -3 : (Pm ? -2 : 2));
  }
  else
    d.code = (efcmp(P.ex, P.fx, P.ey, P.fy) < 0) ? -1 : 1;
  return d;
}

// Real rotation of the register pair (djrot_ with a real tangent; for
// complex data, a rotation with sat == 0 acts on the real and imaginary
// parts alike): (x, y) := (x + t y, y - t x) c, the two results exchanged
// when PERM, the scaling omitted when C1 (c == 1).  Returns the lane's max
// |updated element|.
template <typename T, bool CPLX, int E, bool PERM, bool C1>
__device__ __forceinline__ T rot_real(T (&xr)[E], T (&yr)[E], T (&xi)[E], T (&yi)[E], const T c, const T at, const int nE,
                                      const int sec)
{
  T tG = (T)0;
#pragma unroll
  for (int j = 0; j < E; ++j) {
    if (j >= nE)
      break;
    const T u = fma_(yr[j], at, xr[j]), v = fma_(-xr[j], at, yr[j]);
    T a = PERM ? v : u, b = PERM ? u : v;
    if constexpr (!C1) { a = cscale(a, c, sec); b = cscale(b, c, sec); }
    xr[j] = a;
    yr[j] = b;
    tG = max1p(tG, max1p(vabs(a), vabs(b)));
    if constexpr (CPLX) {
      const T u2 = fma_(yi[j], at, xi[j]), v2 = fma_(-xi[j], at, yi[j]);
      T a2 = PERM ? v2 : u2, b2 = PERM ? u2 : v2;
      if constexpr (!C1) { a2 = cscale(a2, c, sec); b2 = cscale(b2, c, sec); }
      xi[j] = a2;
      yi[j] = b2;
      tG = max1p(tG, max1p(vabs(a2), vabs(b2)));
    }
  }
  return tG;
}

// In-place update of the pair (dvjsvd.c:460-574): x,y in registers, the V
// columns streamed.  Writes the pair's (e,f) as the reference leaves them
// (swapped for codes -1 and -3; the norms of updated columns are recomputed
// at their next step).  All lanes of the group call this with the same
// decision.  With RED = false the update maximum tG is left per lane (not
// reduced over the group, not checked for finiteness): the caller folds it.
template <typename T, bool CPLX, int B, int E, bool RED = true>
__device__ __forceinline__ PairOut<T, CPLX>
pair_apply(const int n, T (&xr)[E], T (&yr)[E], T (&xi)[E], T (&yi)[E],
           T *__restrict__ Vrp, T *__restrict__ Vrq, T *__restrict__ Vip, T *__restrict__ Viq,
           const PairScal<T> &P, const PairDec<T> &D, const int vtrack, T &eSp, T &fSp, T &eSq, T &fSq,
           double *lds, const int lane, const int nE = E)
{
  PairOut<T, CPLX> out{D.code, D.trans, D.err, (T)0, (T)0};
  if (D.err)
    return out;
  const T ex = P.ex, fx = P.fx, ey = P.ey, fy = P.fy;
  const T c = D.c, at = D.at, sat = D.sat;
  T tG = (T)0, tV = (T)0;
  const int code = D.code;
  if ((code == 2) || (code == -2)) {
    const bool perm = (code < 0);
    // G pair in registers: djrot_/zjrot_ with length +-m; a complex
    // rotation with sat == 0 is two real rotations (zjrot.c:26-35)
    if ((at == (T)0) && (!CPLX || (sat == (T)0))) {
      if ((c == (T)1) && perm) {
#pragma unroll
        for (int j = 0; j < E; ++j) {
          if (j >= nE)
            break;
          T t = xr[j]; xr[j] = yr[j]; yr[j] = t;
          if constexpr (CPLX) { t = xi[j]; xi[j] = yi[j]; yi[j] = t; }
        }
      }
    }
    else if (!CPLX || (sat == (T)0)) {
      // the permutation and the unit cosine select the loop, not the element
      const bool c1 = (c == (T)1);
      if (perm)
        tG = c1 ? rot_real<T, CPLX, E, true, true>(xr, yr, xi, yi, c, at, nE, D.sec)
                : rot_real<T, CPLX, E, true, false>(xr, yr, xi, yi, c, at, nE, D.sec);
      else
        tG = c1 ? rot_real<T, CPLX, E, false, true>(xr, yr, xi, yi, c, at, nE, D.sec)
                : rot_real<T, CPLX, E, false, false>(xr, yr, xi, yi, c, at, nE, D.sec);
    }
    else {
      if constexpr (CPLX) {
        const T cna = -at;
        const bool c1 = (c == (T)1);
#pragma unroll
        for (int j = 0; j < E; ++j) {
          if (j >= nE)
            break;
          const T x_r = xr[j], x_i = xi[j], y_r = yr[j], y_i = yi[j];
          T ar = fma_(y_r, at, fma_(-y_i, sat, x_r));
          T ai = fma_(y_r, sat, fma_(y_i, at, x_i));
          T br = fma_(x_r, cna, fma_(-x_i, sat, y_r));
          T bi = fma_(x_r, sat, fma_(x_i, cna, y_i));
          if (!c1) { ar = cscale(ar, c, D.sec); ai = cscale(ai, c, D.sec); br = cscale(br, c, D.sec); bi = cscale(bi, c, D.sec); }
          tG = xmax(tG, xmax(xmax(vabs(ar), vabs(ai)), xmax(vabs(br), vabs(bi))));
          if (perm) { xr[j] = br; xi[j] = bi; yr[j] = ar; yi[j] = ai; }
          else { xr[j] = ar; xi[j] = ai; yr[j] = br; yi[j] = bi; }
        }
      }
    }
    if constexpr (RED)
      tG = red_max<T, B>(tG, lds);
    // V pair, streamed: same rotation with length +-n
    if ((at == (T)0) && (!CPLX || (sat == (T)0))) {
      if ((c == (T)1) && perm) {
        for (int i = lane; i < n; i += B) {
          T t = Vrp[i]; Vrp[i] = Vrq[i]; Vrq[i] = t;
          if constexpr (CPLX) { t = Vip[i]; Vip[i] = Viq[i]; Viq[i] = t; }
        }
      }
    }
    else {
      const bool c1 = (c == (T)1);
      const T cna = -at;
      for (int i = lane; i < n; i += B) {
        const T x_r = Vrp[i], y_r = Vrq[i];
        if (!CPLX || (sat == (T)0)) {
          T a = perm ? fma_(-x_r, at, y_r) : fma_(y_r, at, x_r);
          T b = perm ? fma_(y_r, at, x_r) : fma_(-x_r, at, y_r);
          if (!c1) { a = cscale(a, c, D.sec); b = cscale(b, c, D.sec); }
          Vrp[i] = a;
          Vrq[i] = b;
          tV = max1p(tV, max1p(vabs(a), vabs(b)));
          if constexpr (CPLX) {
            const T x_i = Vip[i], y_i = Viq[i];
            T a2 = perm ? fma_(-x_i, at, y_i) : fma_(y_i, at, x_i);
            T b2 = perm ? fma_(y_i, at, x_i) : fma_(-x_i, at, y_i);
            if (!c1) { a2 = cscale(a2, c, D.sec); b2 = cscale(b2, c, D.sec); }
            Vip[i] = a2;
            Viq[i] = b2;
            tV = max1p(tV, max1p(vabs(a2), vabs(b2)));
          }
        }
        else if constexpr (CPLX) {
          const T x_i = Vip[i], y_i = Viq[i];
          T ar = fma_(y_r, at, fma_(-y_i, sat, x_r));
          T ai = fma_(y_r, sat, fma_(y_i, at, x_i));
          T br = fma_(x_r, cna, fma_(-x_i, sat, y_r));
          T bi = fma_(x_r, sat, fma_(x_i, cna, y_i));
          if (!c1) { ar = cscale(ar, c, D.sec); ai = cscale(ai, c, D.sec); br = cscale(br, c, D.sec); bi = cscale(bi, c, D.sec); }
          tV = xmax(tV, xmax(xmax(vabs(ar), vabs(ai)), xmax(vabs(br), vabs(bi))));
          if (perm) { Vrp[i] = br; Vip[i] = bi; Vrq[i] = ar; Viq[i] = ai; }
          else { Vrp[i] = ar; Vip[i] = ai; Vrq[i] = br; Viq[i] = bi; }
        }
      }
      if (vtrack)
        tV = red_max<T, B>(tV, lds);
      else
        tV = (T)0;
    }
    if (RED && !isfinite_(tG)) { out.err = CPLX ? -630 : -580; return out; }
    if (!isfinite_(tV)) { out.err = CPLX ? -638 : -588; return out; }
    eSp = ex; fSp = fx; eSq = ey; fSq = fy;
  }
  else if ((code == 3) || (code == -3)) {
    // dgsscl_/zgsscl_ (dgsscl.c:5-85, zgsscl.c:5-119) with t = a21
    const bool perm = (code < 0);
    const T x_e = -ex, y_e = -ey;
    const T a21r = P.a21r, a21i = P.a21i;
    if (!perm) {
      const T fy_x = fy / fx;
      const T tfr = -a21r * fy_x;
      const T tfi = a21i * fy_x;
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
      eSp = ex; fSp = fx; eSq = ey; fSq = fy;
    }
    else {
      const T fx_y = fx / fy;
      const T tfr = -a21r * fx_y;
      const T tfi = -a21i * fx_y;
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
      for (int i = lane; i < n; i += B) {
        T t = Vrp[i]; Vrp[i] = Vrq[i]; Vrq[i] = t;
        if constexpr (CPLX) { t = Vip[i]; Vip[i] = Viq[i]; Viq[i] = t; }
      }
      eSp = ey; fSp = fy; eSq = ex; fSq = fx;
    }
    if constexpr (RED)
      tG = red_max<T, B>(tG, lds);
    if (RED && !isfinite_(tG)) { out.err = CPLX ? -630 : -580; return out; }
  }
  else if (code == -1) {
#pragma unroll
    for (int j = 0; j < E; ++j) {
      if (j >= nE)
        break;
      T t = xr[j]; xr[j] = yr[j]; yr[j] = t;
      if constexpr (CPLX) { t = xi[j]; xi[j] = yi[j]; yi[j] = t; }
    }
    for (int i = lane; i < n; i += B) {
      T t = Vrp[i]; Vrp[i] = Vrq[i]; Vrq[i] = t;
      if constexpr (CPLX) { t = Vip[i]; Vip[i] = Viq[i]; Viq[i] = t; }
    }
    eSp = ey; fSp = fy; eSq = ex; fSq = fx;
  }
  else {
    eSp = ex; fSp = fx; eSq = ey; fSq = fy;
  }
  out.tG = tG;
  out.tV = tV;
  return out;
}

// The V part of pair_apply for a V pair held in registers (row lane + B j,
// j < nEV): the same swaps and rotations with the same fused multiply-adds;
// returns the lane's max |updated element| (the caller reduces it when V
// is tracked).  Nothing happens on a failed decision.
template <typename T, bool CPLX, int E>
__device__ __forceinline__ T v_apply(const PairDec<T> &D, T (&vx)[E], T (&vy)[E], T (&vxi)[E], T (&vyi)[E],
                                     const int nEV)
{
  T tV = (T)0;
  if (D.err)
    return tV;
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
      for (int j = 0; j < E; ++j) {
        if (j >= nEV)
          break;
        const T x_r = vx[j], y_r = vy[j];
        if (!CPLX || (sat == (T)0)) {
          T a = perm ? fma_(-x_r, at, y_r) : fma_(y_r, at, x_r);
          T b = perm ? fma_(y_r, at, x_r) : fma_(-x_r, at, y_r);
          if (!c1) { a = cscale(a, c, D.sec); b = cscale(b, c, D.sec); }
          vx[j] = a;
          vy[j] = b;
          tV = max1p(tV, max1p(vabs(a), vabs(b)));
          if constexpr (CPLX) {
            const T x_i = vxi[j], y_i = vyi[j];
            T a2 = perm ? fma_(-x_i, at, y_i) : fma_(y_i, at, x_i);
            T b2 = perm ? fma_(y_i, at, x_i) : fma_(-x_i, at, y_i);
            if (!c1) { a2 = cscale(a2, c, D.sec); b2 = cscale(b2, c, D.sec); }
            vxi[j] = a2;
            vyi[j] = b2;
            tV = max1p(tV, max1p(vabs(a2), vabs(b2)));
          }
        }
        else if constexpr (CPLX) {
          const T x_i = vxi[j], y_i = vyi[j];
          T ar = fma_(y_r, at, fma_(-y_i, sat, x_r));
          T ai = fma_(y_r, sat, fma_(y_i, at, x_i));
          T br = fma_(x_r, cna, fma_(-x_i, sat, y_r));
          T bi = fma_(x_r, sat, fma_(x_i, cna, y_i));
          if (!c1) { ar = cscale(ar, c, D.sec); ai = cscale(ai, c, D.sec); br = cscale(br, c, D.sec); bi = cscale(bi, c, D.sec); }
          tV = xmax(tV, xmax(xmax(vabs(ar), vabs(ai)), xmax(vabs(br), vabs(bi))));
          if (perm) { vx[j] = br; vxi[j] = bi; vy[j] = ar; vyi[j] = ai; }
          else { vx[j] = ar; vxi[j] = ai; vy[j] = br; vyi[j] = bi; }
        }
      }
    }
  }
  if (swap) {
#pragma unroll
    for (int j = 0; j < E; ++j) {
      T t = vx[j]; vx[j] = vy[j]; vy[j] = t;
      if constexpr (CPLX) { t = vxi[j]; vxi[j] = vyi[j]; vyi[j] = t; }
    }
  }
  return tV;
}

// The three phases back to back, for a group that owns one pair.
template <typename T, bool CPLX, int B, int E>
__device__ __forceinline__ PairOut<T, CPLX>
pair_step(const int m, const int n, T (&xr)[E], T (&yr)[E], T (&xi)[E], T (&yi)[E],
          T *__restrict__ Vrp, T *__restrict__ Vrq, T *__restrict__ Vip, T *__restrict__ Viq,
          T &eSp, T &fSp, T &eSq, T &fSq, const Step<T> &S, double *lds, const int lane, PairDec<T> &D)
{
  (void)m;
  const PairScal<T> P = pair_reduce<T, CPLX, B, E>(xr, yr, xi, yi, lds);
  D = pair_decide<T, CPLX>(P, S.tol, S.gst, S.sec);
  return pair_apply<T, CPLX, B, E>(n, xr, yr, xi, yi, Vrp, Vrq, Vip, Viq, P, D, S.vtrack, eSp, fSp, eSq, fSq, lds, lane);
}

} // namespace vj
