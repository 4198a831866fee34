// Adjoint (backward) of the thin real SVD A = U diag(s) V^T for a batch of
// matrices (no reference counterpart; the complex form is k_csvd_bwd below).  With F_ij = 1/(s_j^2 - s_i^2) for
// i != j (0 on the diagonal), J = U^T dU and K = V^T dV:
//   dA = U [diag(ds) + (F o (J - J^T)) S + S (F o (K - K^T))] V^T
//        + (I - U U^T) dU S^-1 V^T
// (V is square, so the (I - V V^T) term vanishes).  It is evaluated as
//   N  = diag(ds) + (F o (J - J^T)) S + S (F o (K - K^T)) - J S^-1,
//   P  = N V^T,
//   dA = U P + (dU S^-1) V^T,
// so that the only intermediates are the n x n blocks J, K (then N, P):
// nothing of size m x n is staged, and m is not limited.
//
// One workgroup of four waves per matrix.  The four products run on the
// FP64 matrix cores (v_mfma_f64_16x16x4f64; FP32 data: the f32 16x16x4 form)
// in 16 x 16 output tiles spread over the waves.  The operands of U, dU, V
// and dV come straight from global memory in the instruction's lane layout
// (lane l holds A(i0 + l%16, k0 + l/16) and B(k0 + l/16, j0 + l%16); its four
// results are four rows of column j0 + l%16 of D), the rows of a column
// contiguous for U and dU.  J and K, then N, then P live in LDS (ld n16 + 1,
// so that the 16 columns a wave reads fall in different banks); each dA tile
// is transposed through a per-wave LDS tile so that its store is coalesced.
#include "svd_common.hpp"
#include "vecjac.h"

#include <algorithm>

namespace vj {

namespace {

// One 16 x 16 x 4 product D = A B + C.  Lane l supplies A(l%16, l/16) and
// B(l/16, l%16); its result r (r = 0..3) is D(row(l/16, r), l%16), with the
// row map of each form as measured on gfx950: l/16 + 4r for FP64,
// 4(l/16) + r for FP32.
template <typename T> struct Mfma;
template <> struct Mfma<double> {
  typedef double V4 __attribute__((ext_vector_type(4)));
  static __device__ __forceinline__ V4 op(double a, double b, V4 c)
  {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int row(int lk, int r) { return lk + 4 * r; }
};
template <> struct Mfma<float> {
  typedef float V4 __attribute__((ext_vector_type(4)));
  static __device__ __forceinline__ V4 op(float a, float b, V4 c)
  {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int row(int lk, int r) { return 4 * lk + r; }
};

// orders the LDS accesses of the lanes of one wave (no workgroup barrier)
__device__ __forceinline__ void wave_sync()
{
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

constexpr int BWD_WAVES = 4;
constexpr int CH = 8;       // steps of 4 rows whose loads are issued together
constexpr int TILE_LD = 17; // per-wave 16 x 16 transpose tile, padded

} // namespace

// LDS: J (then N) and K (then P), n16 x (n16 + 1) each (in a global
// workspace instead when GW); 1/s (n16); one 16 x 17 tile per wave.  PL
// planes: 1 real, 2 complex (real and imaginary parts).
template <typename T> static size_t bwd_jk_elems(int n16, int pl) { return (size_t)2 * pl * n16 * (n16 + 1); }
template <typename T> static size_t bwd_lds_bytes(int n16, int pl, bool gw)
{
  return ((gw ? 0 : bwd_jk_elems<T>(n16, pl)) + n16 + (size_t)BWD_WAVES * 16 * TILE_LD) * sizeof(T);
}

// Tile (ti, tj) of X^T Y (real) for columns of length len: A(i, k) =
// X(k, ti*16 + i), B(k, j) = Y(k, tj*16 + j), X and Y column-major.
template <typename T>
__device__ __forceinline__ typename Mfma<T>::V4 tile_xty(const T *X, const T *Y, const int len, const int n,
                                                         const int ti, const int tj, const int li, const int lk)
{
  using V4 = typename Mfma<T>::V4;
  const int ca = ti * 16 + li, cb = tj * 16 + li;
  const T *const xa = X + (long long)ca * len, *const yb = Y + (long long)cb * len;
  V4 acc = {0, 0, 0, 0};
  // CH steps of 4 rows at a time: all their loads are issued before the
  // chain of products that consumes them
  for (int k0 = 0; k0 < len; k0 += 4 * CH) {
    T a[CH], bv[CH];
#pragma unroll
    for (int c = 0; c < CH; ++c) {
      const int k = k0 + 4 * c + lk;
      a[c] = (ca < n && k < len) ? xa[k] : (T)0;
      bv[c] = (cb < n && k < len) ? yb[k] : (T)0;
    }
#pragma unroll
    for (int c = 0; c < CH; ++c)
      acc = Mfma<T>::op(a[c], bv[c], acc);
  }
  return acc;
}

// The same tile of X^H Y for complex X = Xr + i Xi, Y = Yr + i Yi:
// re = Xr^T Yr + Xi^T Yi, im = Xr^T Yi - Xi^T Yr.
template <typename T>
__device__ __forceinline__ void tile_xhy(const T *Xr, const T *Xi, const T *Yr, const T *Yi, const int len, const int n,
                                         const int ti, const int tj, const int li, const int lk,
                                         typename Mfma<T>::V4 &re, typename Mfma<T>::V4 &im)
{
  const int ca = ti * 16 + li, cb = tj * 16 + li;
  const long long oa = (long long)ca * len, ob = (long long)cb * len;
  re = {0, 0, 0, 0};
  im = {0, 0, 0, 0};
  for (int k0 = 0; k0 < len; k0 += 4 * (CH / 2)) {
    T ar[CH / 2], ai[CH / 2], br[CH / 2], bi[CH / 2];
#pragma unroll
    for (int c = 0; c < CH / 2; ++c) {
      const int k = k0 + 4 * c + lk;
      const bool ia = (ca < n && k < len), ib = (cb < n && k < len);
      ar[c] = ia ? Xr[oa + k] : (T)0;
      ai[c] = ia ? Xi[oa + k] : (T)0;
      br[c] = ib ? Yr[ob + k] : (T)0;
      bi[c] = ib ? Yi[ob + k] : (T)0;
    }
#pragma unroll
    for (int c = 0; c < CH / 2; ++c) {
      re = Mfma<T>::op(ar[c], br[c], re);
      re = Mfma<T>::op(ai[c], bi[c], re);
      im = Mfma<T>::op(ar[c], bi[c], im);
      im = Mfma<T>::op(-ai[c], br[c], im);
    }
  }
}

// Stores a 16 x 16 result tile (rows r0.., column c0..) of an m x n
// column-major output through the wave's LDS tile, so that each column's 16
// rows go out as one contiguous run.
template <typename T>
__device__ __forceinline__ void store_tile(T *out, const int m, const int n, const int r0, const int c0,
                                           const typename Mfma<T>::V4 &acc, T *sT, const int lane)
{
  const int li = lane & 15, lk = lane >> 4;
#pragma unroll
  for (int r = 0; r < 4; ++r)
    sT[li * TILE_LD + Mfma<T>::row(lk, r)] = acc[r];
  wave_sync();
  const int cc = lane >> 2, rr = 4 * (lane & 3);
  const int col = c0 + cc;
  if (col < n) {
    T *const o = out + (long long)col * m + r0 + rr;
#pragma unroll
    for (int q = 0; q < 4; ++q)
      if (r0 + rr + q < m)
        o[q] = sT[cc * TILE_LD + rr + q];
  }
  wave_sync();
}

template <typename T, bool GW>
__global__ __launch_bounds__(64 * BWD_WAVES) void k_svd_bwd(const int nb, const int m, const int n, const int n16,
                                                             const T *__restrict__ U, const T *__restrict__ s,
                                                             const T *__restrict__ V, const T *__restrict__ dU,
                                                             const T *__restrict__ ds, const T *__restrict__ dV,
                                                             T *__restrict__ dA, T *__restrict__ ws)
{
  // per matrix b: U, dU, dA are m x n column-major (ld m) at b*m*n; V, dV are
  // n x n (ld n) at b*n*n; s, ds at b*n.
  using V4 = typename Mfma<T>::V4;
  extern __shared__ double smem_[];
  const int ld = n16 + 1;
  T *const sJ = GW ? ws + (long long)blockIdx.x * 2 * n16 * ld : (T *)smem_; // J, then N
  T *const sK = sJ + n16 * ld;                                               // K, then P
  T *const sR = GW ? (T *)smem_ : sK + n16 * ld;                            // 1/s_k (0 past n)
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  T *const sT = sR + n16 + w * 16 * TILE_LD;
  const int li = lane & 15, lk = lane >> 4;
  const int nt = n16 >> 4, mt = (m + 15) >> 4;
  for (int b = blockIdx.x; b < nb; b += gridDim.x) {
    const T *const Ub = U + (long long)b * m * n, *const dUb = dU + (long long)b * m * n;
    const T *const Vb = V + (long long)b * n * n, *const dVb = dV + (long long)b * n * n;
    const T *const sb = s + (long long)b * n, *const dsb = ds + (long long)b * n;
    for (int t = threadIdx.x; t < n16; t += 64 * BWD_WAVES)
      sR[t] = (t < n) ? (T)1 / sb[t] : (T)0;
    // J = U^T dU (inner m) and K = V^T dV (inner n)
    for (int task = w; task < 2 * nt * nt; task += BWD_WAVES) {
      const bool isK = task >= nt * nt;
      const int tt = isK ? task - nt * nt : task;
      const int ti = tt % nt, tj = tt / nt;
      const V4 acc = isK ? tile_xty<T>(Vb, dVb, n, n, ti, tj, li, lk) : tile_xty<T>(Ub, dUb, m, n, ti, tj, li, lk);
      T *const dst = isK ? sK : sJ;
      const int cb = tj * 16 + li;
#pragma unroll
      for (int r = 0; r < 4; ++r)
        dst[cb * ld + ti * 16 + Mfma<T>::row(lk, r)] = acc[r];
    }
    __syncthreads();
    // N in place of J.  One thread owns the unordered pair {i, j}: it reads
    // J(i,j), J(j,i), K(i,j), K(j,i) and writes N(i,j), N(j,i), so no other
    // thread touches those words.  Outside n x n, N = 0.
    for (int t = threadIdx.x; t < n16 * n16; t += 64 * BWD_WAVES) {
      const int i = t % n16, j = t / n16;
      if (i > j)
        continue;
      if (j >= n) {
        sJ[j * ld + i] = (T)0;
        sJ[i * ld + j] = (T)0;
        continue;
      }
      if (i == j) {
        sJ[i * ld + i] = dsb[i] - sJ[i * ld + i] * sR[i];
        continue;
      }
      const T si = sb[i], sj = sb[j];
      const T F = (T)1 / ((sj - si) * (sj + si));
      const T Jij = sJ[j * ld + i], Jji = sJ[i * ld + j];
      const T aJ = Jij - Jji;
      const T aK = sK[j * ld + i] - sK[i * ld + j];
      sJ[j * ld + i] = F * (aJ * sj + si * aK) - Jij * sR[j];
      sJ[i * ld + j] = F * (aJ * si + sj * aK) - Jji * sR[i];
    }
    __syncthreads();
    // P = N V^T in place of K: A(i, k) = N(ti*16 + i, k), B(k, c) =
    // V(tj*16 + c, k) = V[k n + c], the 16 entries of a row contiguous
    for (int tt = w; tt < nt * nt; tt += BWD_WAVES) {
      const int ti = tt % nt, tj = tt / nt;
      const int ia = ti * 16 + li, cb = tj * 16 + li;
      V4 acc = {0, 0, 0, 0};
      for (int k0 = 0; k0 < n16; k0 += 16) {
        T a[4], bv[4];
#pragma unroll
        for (int c = 0; c < 4; ++c) {
          const int k = k0 + 4 * c + lk;
          a[c] = sJ[k * ld + ia];
          bv[c] = (k < n && cb < n) ? Vb[(long long)k * n + cb] : (T)0;
        }
#pragma unroll
        for (int c = 0; c < 4; ++c)
          acc = Mfma<T>::op(a[c], bv[c], acc);
      }
#pragma unroll
      for (int r = 0; r < 4; ++r)
        sK[cb * ld + ti * 16 + Mfma<T>::row(lk, r)] = acc[r];
    }
    __syncthreads();
    // dA = U P + (dU S^-1) V^T, tile (ri, cj): A(r, k) = U(ri*16 + r, k) or
    // dU(ri*16 + r, k) / s_k, B(k, c) = P(k, cj*16 + c) or V(cj*16 + c, k)
    for (int tt = w; tt < mt * nt; tt += BWD_WAVES) {
      const int ri = tt % mt, cj = tt / mt;
      const int ra = ri * 16 + li, cb = cj * 16 + li;
      const bool rin = ra < m;
      V4 acc = {0, 0, 0, 0};
      for (int k0 = 0; k0 < n16; k0 += 16) {
        T a1[4], b1[4], a2[4], b2[4];
#pragma unroll
        for (int c = 0; c < 4; ++c) {
          const int k = k0 + 4 * c + lk;
          const bool kin = k < n;
          a1[c] = (rin && kin) ? Ub[(long long)k * m + ra] : (T)0;
          a2[c] = (rin && kin) ? dUb[(long long)k * m + ra] : (T)0;
          b2[c] = (kin && cb < n) ? Vb[(long long)k * n + cb] : (T)0;
          b1[c] = sK[cb * ld + k];
        }
#pragma unroll
        for (int c = 0; c < 4; ++c) {
          acc = Mfma<T>::op(a1[c], b1[c], acc);
          acc = Mfma<T>::op(a2[c] * sR[k0 + 4 * c + lk], b2[c], acc);
        }
      }
      store_tile<T>(dA + (long long)b * m * n, m, n, ri * 16, cj * 16, acc, sT, lane);
    }
    __syncthreads();
  }
}

// Staged form for small matrices (every operand of one matrix in LDS): the
// same four products and N as k_svd_bwd, with U, dU, V, dV, s and ds copied
// into LDS once per matrix (the columns of U, dU, V and dV padded to an odd
// leading dimension, so that the 16 columns an MFMA operand reads fall in
// different banks) and the next matrix's copy loaded into registers while
// this one is computed, so that the products wait on LDS, never on device
// memory.  The two tiles a wave owns in a phase run as two independent
// chains of MFMAs, interleaved.
constexpr int STG_U = 8;  // staged elements per lane of U and of dU (m n <= 2048)
constexpr int STG_V = 8;  // of V and of dV (n n <= 2048)
constexpr int STG_LANES = 64 * BWD_WAVES;
constexpr int STG_CH = 4; // steps of 4 rows whose LDS operands are read together

struct StgLayout {
  int ldU, ldV;             // padded leading dimensions of the staged U, dU and V, dV
  size_t oU, odU, oV, odV;  // offsets (elements) past the k_svd_bwd part of LDS
  size_t oS, odS, total;
};

__host__ __device__ inline StgLayout stg_layout(int m, int n)
{
  StgLayout L;
  L.ldU = m | 1;
  L.ldV = n | 1;
  L.oU = 0;
  L.odU = L.oU + (size_t)n * L.ldU;
  L.oV = L.odU + (size_t)n * L.ldU;
  L.odV = L.oV + (size_t)n * L.ldV;
  L.oS = L.odV + (size_t)n * L.ldV;
  L.odS = L.oS + (size_t)n;
  L.total = L.odS + (size_t)n;
  return L;
}

template <typename T> static bool stg_fits(int m, int n)
{
  return (n <= 64) && ((long long)m * n <= (long long)STG_LANES * STG_U) && (n * n <= STG_LANES * STG_V) &&
         (bwd_lds_bytes<T>((n + 15) & ~15, 1, false) + stg_layout(m, n).total * sizeof(T) <= 160 * 1024);
}

// the interleaved pair of J / K tiles (X^T Y of columns of length len, X
// and Y staged with leading dimension ldX) of tasks task and task2 (< 0:
// none)
template <typename T>
__device__ __forceinline__ void stg_tile2(const T *X, const T *Y, const T *X2, const T *Y2, const int ldX,
                                          const int ldX2, const int len, const int len2, const int n, const int ti,
                                          const int tj, const int ti2, const int tj2, const bool two, const int li,
                                          const int lk, typename Mfma<T>::V4 &acc, typename Mfma<T>::V4 &acc2)
{
  const int ca = ti * 16 + li, cb = tj * 16 + li, ca2 = ti2 * 16 + li, cb2 = tj2 * 16 + li;
  const int lmax = (two && len2 > len) ? len2 : len;
  acc = {0, 0, 0, 0};
  acc2 = {0, 0, 0, 0};
  for (int k0 = 0; k0 < lmax; k0 += 4 * STG_CH) {
    T a[STG_CH], bv[STG_CH], a2[STG_CH], b2[STG_CH];
#pragma unroll
    for (int c = 0; c < STG_CH; ++c) {
      const int k = k0 + 4 * c + lk;
      a[c] = (ca < n && k < len) ? X[ca * ldX + k] : (T)0;
      bv[c] = (cb < n && k < len) ? Y[cb * ldX + k] : (T)0;
      a2[c] = (two && ca2 < n && k < len2) ? X2[ca2 * ldX2 + k] : (T)0;
      b2[c] = (two && cb2 < n && k < len2) ? Y2[cb2 * ldX2 + k] : (T)0;
    }
#pragma unroll
    for (int c = 0; c < STG_CH; ++c) {
      if (k0 + 4 * c < len)
        acc = Mfma<T>::op(a[c], bv[c], acc);
      if (two && (k0 + 4 * c < len2))
        acc2 = Mfma<T>::op(a2[c], b2[c], acc2);
    }
  }
}

template <typename T>
__global__ __launch_bounds__(STG_LANES, 2) void k_svd_bwd_stg(const int nb, const int m, const int n, const int n16,
                                                           const T *__restrict__ U, const T *__restrict__ s,
                                                           const T *__restrict__ V, const T *__restrict__ dU,
                                                           const T *__restrict__ ds, const T *__restrict__ dV,
                                                           T *__restrict__ dA)
{
  using V4 = typename Mfma<T>::V4;
  extern __shared__ double smem_[];
  const int ld = n16 + 1;
  T *const sJ = (T *)smem_;           // J, then N
  T *const sK = sJ + n16 * ld;        // K, then P
  T *const sR = sK + n16 * ld;        // 1/s_k (0 past n)
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  T *const sT = sR + n16 + w * 16 * TILE_LD;
  T *const stg = sR + n16 + BWD_WAVES * 16 * TILE_LD;
  const StgLayout L = stg_layout(m, n);
  T *const sU = stg + L.oU, *const sdU = stg + L.odU, *const sV = stg + L.oV, *const sdV = stg + L.odV;
  T *const sS = stg + L.oS, *const sdS = stg + L.odS;
  const int li = lane & 15, lk = lane >> 4;
  const int nt = n16 >> 4, mt = (m + 15) >> 4;
  const int mn = m * n, nn = n * n;
  // element t + STG_LANES j of a column-major array of R rows sits in row
  // (t + STG_LANES j) % R of column (t + STG_LANES j) / R: the staged copy
  // walks it by steps of STG_LANES
  const int iU0 = t % m, kU0 = t / m, dIU = STG_LANES % m, dKU = STG_LANES / m;
  const int iV0 = t % n, kV0 = t / n, dIV = STG_LANES % n, dKV = STG_LANES / n;
  T pU[STG_U], pdU[STG_U], pV[STG_V], pdV[STG_V], ps = (T)0, pds = (T)0;
  auto fetch = [&](const int b) {
    const T *const Ub = U + (long long)b * mn, *const dUb = dU + (long long)b * mn;
    const T *const Vb = V + (long long)b * nn, *const dVb = dV + (long long)b * nn;
#pragma unroll
    for (int j = 0; j < STG_U; ++j) {
      const int e = t + STG_LANES * j;
      pU[j] = (e < mn) ? Ub[e] : (T)0;
      pdU[j] = (e < mn) ? dUb[e] : (T)0;
    }
#pragma unroll
    for (int j = 0; j < STG_V; ++j) {
      const int e = t + STG_LANES * j;
      pV[j] = (e < nn) ? Vb[e] : (T)0;
      pdV[j] = (e < nn) ? dVb[e] : (T)0;
    }
    ps = (t < n) ? s[(long long)b * n + t] : (T)0;
    pds = (t < n) ? ds[(long long)b * n + t] : (T)0;
  };
  auto stage = [&]() {
    int i = iU0, k = kU0;
#pragma unroll
    for (int j = 0; j < STG_U; ++j) {
      if (t + STG_LANES * j < mn) {
        sU[k * L.ldU + i] = pU[j];
        sdU[k * L.ldU + i] = pdU[j];
      }
      i += dIU;
      k += dKU;
      if (i >= m) {
        i -= m;
        ++k;
      }
    }
    i = iV0;
    k = kV0;
#pragma unroll
    for (int j = 0; j < STG_V; ++j) {
      if (t + STG_LANES * j < nn) {
        sV[k * L.ldV + i] = pV[j];
        sdV[k * L.ldV + i] = pdV[j];
      }
      i += dIV;
      k += dKV;
      if (i >= n) {
        i -= n;
        ++k;
      }
    }
    if (t < n) {
      sS[t] = ps;
      sdS[t] = pds;
    }
  };
  int b = blockIdx.x;
  if (b >= nb)
    return;
  fetch(b);
  stage();
  __syncthreads();
  for (; b < nb; b += gridDim.x) {
    const int bn = b + gridDim.x;
    if (bn < nb)
      fetch(bn); // the next matrix in flight while this one is computed
    for (int q = t; q < n16; q += STG_LANES)
      sR[q] = (q < n) ? (T)1 / sS[q] : (T)0;
    // J = U^T dU (inner m) and K = V^T dV (inner n), two tiles at a time
    const int ntask = 2 * nt * nt;
    for (int task = w; task < ntask; task += 2 * BWD_WAVES) {
      const int task2 = task + BWD_WAVES;
      const bool two = task2 < ntask;
      const bool isK = task >= nt * nt, isK2 = task2 >= nt * nt;
      const int tt = isK ? task - nt * nt : task, tt2 = isK2 ? task2 - nt * nt : task2;
      const int ti = tt % nt, tj = tt / nt, ti2 = tt2 % nt, tj2 = tt2 / nt;
      V4 acc, acc2;
      stg_tile2<T>(isK ? sV : sU, isK ? sdV : sdU, isK2 ? sV : sU, isK2 ? sdV : sdU, isK ? L.ldV : L.ldU,
                   isK2 ? L.ldV : L.ldU, isK ? n : m, isK2 ? n : m, n, ti, tj, ti2, tj2, two, li, lk, acc, acc2);
      {
        T *const dst = isK ? sK : sJ;
        const int cb = tj * 16 + li;
#pragma unroll
        for (int r = 0; r < 4; ++r)
          dst[cb * ld + ti * 16 + Mfma<T>::row(lk, r)] = acc[r];
      }
      if (two) {
        T *const dst = isK2 ? sK : sJ;
        const int cb = tj2 * 16 + li;
#pragma unroll
        for (int r = 0; r < 4; ++r)
          dst[cb * ld + ti2 * 16 + Mfma<T>::row(lk, r)] = acc2[r];
      }
    }
    __syncthreads();
    // N in place of J, as in k_svd_bwd
    for (int q = t; q < n16 * n16; q += STG_LANES) {
      const int i = q % n16, j = q / n16;
      if (i > j)
        continue;
      if (j >= n) {
        sJ[j * ld + i] = (T)0;
        sJ[i * ld + j] = (T)0;
        continue;
      }
      if (i == j) {
        sJ[i * ld + i] = sdS[i] - sJ[i * ld + i] * sR[i];
        continue;
      }
      const T si = sS[i], sj = sS[j];
      const T F = (T)1 / ((sj - si) * (sj + si));
      const T Jij = sJ[j * ld + i], Jji = sJ[i * ld + j];
      const T aJ = Jij - Jji;
      const T aK = sK[j * ld + i] - sK[i * ld + j];
      sJ[j * ld + i] = F * (aJ * sj + si * aK) - Jij * sR[j];
      sJ[i * ld + j] = F * (aJ * si + sj * aK) - Jji * sR[i];
    }
    __syncthreads();
    // P = N V^T in place of K
    for (int tt = w; tt < nt * nt; tt += BWD_WAVES) {
      const int ti = tt % nt, tj = tt / nt;
      const int ia = ti * 16 + li, cb = tj * 16 + li;
      V4 acc = {0, 0, 0, 0};
      for (int k0 = 0; k0 < n16; k0 += 16) {
        T a[4], bv[4];
#pragma unroll
        for (int c = 0; c < 4; ++c) {
          const int k = k0 + 4 * c + lk;
          a[c] = sJ[k * ld + ia];
          bv[c] = (k < n && cb < n) ? sV[k * L.ldV + cb] : (T)0;
        }
#pragma unroll
        for (int c = 0; c < 4; ++c)
          acc = Mfma<T>::op(a[c], bv[c], acc);
      }
#pragma unroll
      for (int r = 0; r < 4; ++r)
        sK[cb * ld + ti * 16 + Mfma<T>::row(lk, r)] = acc[r];
    }
    __syncthreads();
    // dA = U P + (dU S^-1) V^T
    for (int tt = w; tt < mt * nt; tt += BWD_WAVES) {
      const int ri = tt % mt, cj = tt / mt;
      const int ra = ri * 16 + li, cb = cj * 16 + li;
      const bool rin = ra < m;
      V4 acc = {0, 0, 0, 0};
      for (int k0 = 0; k0 < n16; k0 += 16) {
        T a1[4], b1[4], a2[4], b2[4];
#pragma unroll
        for (int c = 0; c < 4; ++c) {
          const int k = k0 + 4 * c + lk;
          const bool kin = k < n;
          a1[c] = (rin && kin) ? sU[k * L.ldU + ra] : (T)0;
          a2[c] = (rin && kin) ? sdU[k * L.ldU + ra] * sR[k] : (T)0;
          b2[c] = (kin && cb < n) ? sV[k * L.ldV + cb] : (T)0;
          b1[c] = sK[cb * ld + k];
        }
#pragma unroll
        for (int c = 0; c < 4; ++c) {
          acc = Mfma<T>::op(a1[c], b1[c], acc);
          acc = Mfma<T>::op(a2[c], b2[c], acc);
        }
      }
      store_tile<T>(dA + (long long)b * m * n, m, n, ri * 16, cj * 16, acc, sT, lane);
    }
    __syncthreads();
    if (bn < nb) {
      stage();
      __syncthreads();
    }
  }
}

// Complex adjoint (split real and imaginary planes), for a loss that does
// not depend on the phase of the singular vector pairs (u_k, v_k) ->
// (u_k e^{i phi}, v_k e^{i phi}).  With J = U^H dU, K = V^H dV:
//   N_ij = F_ij ((J - J^H)_ij s_j + s_i (K - K^H)_ij) - J_ij / s_j  (i != j),
//   N_ii = ds_i - J_ii / s_i + i (Im J_ii - Im K_ii) / (2 s_i),
//   P = N V^H,  dA = U P + (dU S^-1) V^H,
// each complex product as four real ones on the matrix cores (a difference
// by a negated operand, which is exact).  The diagonal phase term is the one
// the gauge leaves: only the difference of the phase changes of u_k and v_k
// is determined by dA, and for a phase-independent loss Im K_ii = -Im J_ii.
template <typename T, bool GW>
__global__ __launch_bounds__(64 * BWD_WAVES) void k_csvd_bwd(const int nb, const int m, const int n, const int n16,
                                                              const T *__restrict__ Ur, const T *__restrict__ Ui,
                                                              const T *__restrict__ s, const T *__restrict__ Vr,
                                                              const T *__restrict__ Vi, const T *__restrict__ dUr,
                                                              const T *__restrict__ dUi, const T *__restrict__ ds,
                                                              const T *__restrict__ dVr, const T *__restrict__ dVi,
                                                              T *__restrict__ dAr, T *__restrict__ dAi,
                                                              T *__restrict__ ws)
{
  using V4 = typename Mfma<T>::V4;
  extern __shared__ double smem_[];
  const int ld = n16 + 1, sz = n16 * ld;
  T *const sJr = GW ? ws + (long long)blockIdx.x * 4 * sz : (T *)smem_; // J, then N
  T *const sJi = sJr + sz;
  T *const sKr = sJi + sz;                                              // K, then P
  T *const sKi = sKr + sz;
  T *const sR = GW ? (T *)smem_ : sKi + sz;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  T *const sT = sR + n16 + w * 16 * TILE_LD;
  const int li = lane & 15, lk = lane >> 4;
  const int nt = n16 >> 4, mt = (m + 15) >> 4;
  for (int b = blockIdx.x; b < nb; b += gridDim.x) {
    const long long oU = (long long)b * m * n, oV = (long long)b * n * n;
    const T *const sb = s + (long long)b * n, *const dsb = ds + (long long)b * n;
    for (int t = threadIdx.x; t < n16; t += 64 * BWD_WAVES)
      sR[t] = (t < n) ? (T)1 / sb[t] : (T)0;
    for (int task = w; task < 2 * nt * nt; task += BWD_WAVES) {
      const bool isK = task >= nt * nt;
      const int tt = isK ? task - nt * nt : task;
      const int ti = tt % nt, tj = tt / nt;
      V4 re, im;
      if (isK)
        tile_xhy<T>(Vr + oV, Vi + oV, dVr + oV, dVi + oV, n, n, ti, tj, li, lk, re, im);
      else
        tile_xhy<T>(Ur + oU, Ui + oU, dUr + oU, dUi + oU, m, n, ti, tj, li, lk, re, im);
      T *const dr = isK ? sKr : sJr, *const di = isK ? sKi : sJi;
      const int cb = tj * 16 + li;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        dr[cb * ld + ti * 16 + Mfma<T>::row(lk, r)] = re[r];
        di[cb * ld + ti * 16 + Mfma<T>::row(lk, r)] = im[r];
      }
    }
    __syncthreads();
    for (int t = threadIdx.x; t < n16 * n16; t += 64 * BWD_WAVES) {
      const int i = t % n16, j = t / n16;
      if (i > j)
        continue;
      const int ij = j * ld + i, ji = i * ld + j;
      if (j >= n) {
        sJr[ij] = sJi[ij] = sJr[ji] = sJi[ji] = (T)0;
        continue;
      }
      if (i == j) {
        const T r = sR[i];
        const T jr = sJr[ij], ji_ = sJi[ij];
        sJr[ij] = dsb[i] - jr * r;
        sJi[ij] = (T)0.5 * (ji_ - sKi[ij]) * r - ji_ * r;
        continue;
      }
      const T si = sb[i], sj = sb[j];
      const T F = (T)1 / ((sj - si) * (sj + si));
      const T Jr_ij = sJr[ij], Ji_ij = sJi[ij], Jr_ji = sJr[ji], Ji_ji = sJi[ji];
      // (J - J^H)_ij = (Jr_ij - Jr_ji) + i (Ji_ij + Ji_ji); (J - J^H)_ji is its
      // negated conjugate and F_ji = -F_ij, so N_ji takes F (ajr s_i + s_j akr)
      // as real part and -F (aji s_i + s_j aki) as imaginary part
      const T ajr = Jr_ij - Jr_ji, aji = Ji_ij + Ji_ji;
      const T akr = sKr[ij] - sKr[ji], aki = sKi[ij] + sKi[ji];
      sJr[ij] = F * (ajr * sj + si * akr) - Jr_ij * sR[j];
      sJi[ij] = F * (aji * sj + si * aki) - Ji_ij * sR[j];
      sJr[ji] = F * (ajr * si + sj * akr) - Jr_ji * sR[i];
      sJi[ji] = -F * (aji * si + sj * aki) - Ji_ji * sR[i];
    }
    __syncthreads();
    // P = N V^H in place of K: B(k, c) = conj(V(tj*16 + c, k)), so
    // Pr = Nr Vr^T + Ni Vi^T and Pi = Ni Vr^T - Nr Vi^T
    for (int tt = w; tt < nt * nt; tt += BWD_WAVES) {
      const int ti = tt % nt, tj = tt / nt;
      const int ia = ti * 16 + li, cb = tj * 16 + li;
      V4 pr = {0, 0, 0, 0}, pi = {0, 0, 0, 0};
      for (int k0 = 0; k0 < n16; k0 += 16) {
        T ar[4], ai[4], br[4], bi[4];
#pragma unroll
        for (int c = 0; c < 4; ++c) {
          const int k = k0 + 4 * c + lk;
          const bool in = (k < n && cb < n);
          ar[c] = sJr[k * ld + ia];
          ai[c] = sJi[k * ld + ia];
          br[c] = in ? Vr[oV + (long long)k * n + cb] : (T)0;
          bi[c] = in ? Vi[oV + (long long)k * n + cb] : (T)0;
        }
#pragma unroll
        for (int c = 0; c < 4; ++c) {
          pr = Mfma<T>::op(ar[c], br[c], pr);
          pr = Mfma<T>::op(ai[c], bi[c], pr);
          pi = Mfma<T>::op(ai[c], br[c], pi);
          pi = Mfma<T>::op(-ar[c], bi[c], pi);
        }
      }
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        sKr[cb * ld + ti * 16 + Mfma<T>::row(lk, r)] = pr[r];
        sKi[cb * ld + ti * 16 + Mfma<T>::row(lk, r)] = pi[r];
      }
    }
    __syncthreads();
    // dA = U P + (dU S^-1) V^H:
    //   re = Ur Pr - Ui Pi + (dUr S^-1) Vr^T + (dUi S^-1) Vi^T
    //   im = Ur Pi + Ui Pr + (dUi S^-1) Vr^T - (dUr S^-1) Vi^T
    for (int tt = w; tt < mt * nt; tt += BWD_WAVES) {
      const int ri = tt % mt, cj = tt / mt;
      const int ra = ri * 16 + li, cb = cj * 16 + li;
      const bool rin = ra < m;
      V4 re = {0, 0, 0, 0}, im = {0, 0, 0, 0};
      for (int k0 = 0; k0 < n16; k0 += 16) {
#pragma unroll
        for (int c = 0; c < 4; ++c) {
          const int k = k0 + 4 * c + lk;
          const bool kin = k < n;
          const long long ou = oU + (long long)k * m + ra, ov = oV + (long long)k * n + cb;
          const T ur = (rin && kin) ? Ur[ou] : (T)0, ui = (rin && kin) ? Ui[ou] : (T)0;
          const T rs = sR[k0 + 4 * c + lk];
          const T gr = (rin && kin) ? dUr[ou] * rs : (T)0, gi = (rin && kin) ? dUi[ou] * rs : (T)0;
          const bool vin = (kin && cb < n);
          const T vr = vin ? Vr[ov] : (T)0, vi = vin ? Vi[ov] : (T)0;
          const T pr = sKr[cb * ld + k], pi = sKi[cb * ld + k];
          re = Mfma<T>::op(ur, pr, re);
          re = Mfma<T>::op(-ui, pi, re);
          re = Mfma<T>::op(gr, vr, re);
          re = Mfma<T>::op(gi, vi, re);
          im = Mfma<T>::op(ur, pi, im);
          im = Mfma<T>::op(ui, pr, im);
          im = Mfma<T>::op(gi, vr, im);
          im = Mfma<T>::op(-gr, vi, im);
        }
      }
      store_tile<T>(dAr + oU, m, n, ri * 16, cj * 16, re, sT, lane);
      store_tile<T>(dAi + oU, m, n, ri * 16, cj * 16, im, sT, lane);
    }
    __syncthreads();
  }
}

// LDS-resident intermediates when they fit a CU's 160 KiB, else a global
// workspace of one J/K block per workgroup (stream-ordered allocation,
// at most two workgroups per CU).
struct BwdLaunch {
  int grid;
  size_t lds;
  bool gw;
};

template <typename T> static int bwd_shape(int nb, int n16, int pl, BwdLaunch &L)
{
  int dev = 0, cus = 0;
  VJ_HIP(hipGetDevice(&dev));
  VJ_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
  L.gw = bwd_lds_bytes<T>(n16, pl, false) > 160 * 1024;
  L.lds = bwd_lds_bytes<T>(n16, pl, L.gw);
  const int per_cu = L.gw ? 2 : (int)std::max<size_t>(1, std::min<size_t>(8, (160 * 1024) / L.lds));
  L.grid = nb < per_cu * cus ? nb : per_cu * cus;
  return 0;
}

template <typename T>
int svd_bwd(int nb, int m, int n, const T *U, const T *s, const T *V, const T *dU, const T *ds, const T *dV, T *dA,
            hipStream_t st)
{
  if (nb <= 0)
    return 0;
  if (n < 1 || m < n)
    return set_error(-2, "svd backward: need 1 <= n <= m");
  const int n16 = (n + 15) & ~15;
  if (stg_fits<T>(m, n)) {
    // small matrices: every operand staged in LDS
    int dev = 0, cus = 0;
    VJ_HIP(hipGetDevice(&dev));
    VJ_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    const size_t lds = bwd_lds_bytes<T>(n16, 1, false) + stg_layout(m, n).total * sizeof(T);
    const int per_cu = (int)std::max<size_t>(1, (160 * 1024) / lds);
    const int grid = nb < per_cu * cus ? nb : per_cu * cus;
    if (lds > 64 * 1024)
      VJ_HIP(hipFuncSetAttribute((const void *)k_svd_bwd_stg<T>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    hipLaunchKernelGGL((k_svd_bwd_stg<T>), dim3(grid), dim3(STG_LANES), lds, st, nb, m, n, n16, U, s, V, dU, ds, dV,
                       dA);
    return check_launch();
  }
  BwdLaunch L{};
  if (int rc = bwd_shape<T>(nb, n16, 1, L))
    return rc;
  T *ws = nullptr;
  if (L.gw) {
    VJ_HIP(hipMallocAsync((void **)&ws, (size_t)L.grid * bwd_jk_elems<T>(n16, 1) * sizeof(T), st));
    hipLaunchKernelGGL((k_svd_bwd<T, true>), dim3(L.grid), dim3(64 * BWD_WAVES), L.lds, st, nb, m, n, n16, U, s, V,
                       dU, ds, dV, dA, ws);
    const int rc = check_launch();
    VJ_HIP(hipFreeAsync(ws, st));
    return rc;
  }
  if (L.lds > 64 * 1024)
    VJ_HIP(hipFuncSetAttribute((const void *)k_svd_bwd<T, false>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)L.lds));
  hipLaunchKernelGGL((k_svd_bwd<T, false>), dim3(L.grid), dim3(64 * BWD_WAVES), L.lds, st, nb, m, n, n16, U, s, V, dU,
                     ds, dV, dA, ws);
  return check_launch();
}

template <typename T>
int csvd_bwd(int nb, int m, int n, const T *Ur, const T *Ui, const T *s, const T *Vr, const T *Vi, const T *dUr,
             const T *dUi, const T *ds, const T *dVr, const T *dVi, T *dAr, T *dAi, hipStream_t st)
{
  if (nb <= 0)
    return 0;
  if (n < 1 || m < n)
    return set_error(-2, "svd backward: need 1 <= n <= m");
  const int n16 = (n + 15) & ~15;
  BwdLaunch L{};
  if (int rc = bwd_shape<T>(nb, n16, 2, L))
    return rc;
  T *ws = nullptr;
  if (L.gw) {
    VJ_HIP(hipMallocAsync((void **)&ws, (size_t)L.grid * bwd_jk_elems<T>(n16, 2) * sizeof(T), st));
    hipLaunchKernelGGL((k_csvd_bwd<T, true>), dim3(L.grid), dim3(64 * BWD_WAVES), L.lds, st, nb, m, n, n16, Ur, Ui, s,
                       Vr, Vi, dUr, dUi, ds, dVr, dVi, dAr, dAi, ws);
    const int rc = check_launch();
    VJ_HIP(hipFreeAsync(ws, st));
    return rc;
  }
  if (L.lds > 64 * 1024)
    VJ_HIP(hipFuncSetAttribute((const void *)k_csvd_bwd<T, false>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)L.lds));
  hipLaunchKernelGGL((k_csvd_bwd<T, false>), dim3(L.grid), dim3(64 * BWD_WAVES), L.lds, st, nb, m, n, n16, Ur, Ui, s,
                     Vr, Vi, dUr, dUi, ds, dVr, dVi, dAr, dAi, ws);
  return check_launch();
}

template int svd_bwd<double>(int, int, int, const double *, const double *, const double *, const double *, const double *,
                             const double *, double *, hipStream_t);
template int svd_bwd<float>(int, int, int, const float *, const float *, const float *, const float *, const float *,
                            const float *, float *, hipStream_t);

} // namespace vj

extern "C" {

// dA for a batch of nb thin SVDs of m x n matrices; see the header.
vjh_int vjh_dsvd_bwd(vjh_int nb, vjh_int m, vjh_int n, const double *U, const double *s, const double *V,
                     const double *dU, const double *ds, const double *dV, double *dA, void *stream)
{
  return vj::svd_bwd<double>((int)nb, (int)m, (int)n, U, s, V, dU, ds, dV, dA, (hipStream_t)stream);
}
vjh_int vjh_ssvd_bwd(vjh_int nb, vjh_int m, vjh_int n, const float *U, const float *s, const float *V,
                     const float *dU, const float *ds, const float *dV, float *dA, void *stream)
{
  return vj::svd_bwd<float>((int)nb, (int)m, (int)n, U, s, V, dU, ds, dV, dA, (hipStream_t)stream);
}
vjh_int vjh_zsvd_bwd(vjh_int nb, vjh_int m, vjh_int n, const double *Ur, const double *Ui, const double *s,
                     const double *Vr, const double *Vi, const double *dUr, const double *dUi, const double *ds,
                     const double *dVr, const double *dVi, double *dAr, double *dAi, void *stream)
{
  return vj::csvd_bwd<double>((int)nb, (int)m, (int)n, Ur, Ui, s, Vr, Vi, dUr, dUi, ds, dVr, dVi, dAr, dAi,
                              (hipStream_t)stream);
}
vjh_int vjh_csvd_bwd(vjh_int nb, vjh_int m, vjh_int n, const float *Ur, const float *Ui, const float *s,
                     const float *Vr, const float *Vi, const float *dUr, const float *dUi, const float *ds,
                     const float *dVr, const float *dVi, float *dAr, float *dAi, void *stream)
{
  return vj::csvd_bwd<float>((int)nb, (int)m, (int)n, Ur, Ui, s, Vr, Vi, dUr, dUi, ds, dVr, dVi, dAr, dAi,
                             (hipStream_t)stream);
}

} // extern "C"
