// Adjoint (backward) of the thin real SVD A = U diag(s) V^T for a batch of
// matrices (no reference counterpart).  With F_ij = 1/(s_j^2 - s_i^2) for
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

// LDS: J (then N) and K (then P), n16 x (n16 + 1) each; 1/s (n16); one
// 16 x 17 tile per wave.
template <typename T> static size_t bwd_lds_bytes(int n16)
{
  return ((size_t)2 * n16 * (n16 + 1) + n16 + (size_t)BWD_WAVES * 16 * TILE_LD) * sizeof(T);
}

template <typename T>
__global__ __launch_bounds__(64 * BWD_WAVES) void k_svd_bwd(const int nb, const int m, const int n, const int n16,
                                                             const T *__restrict__ U, const T *__restrict__ s,
                                                             const T *__restrict__ V, const T *__restrict__ dU,
                                                             const T *__restrict__ ds, const T *__restrict__ dV,
                                                             T *__restrict__ dA)
{
  // per matrix b: U, dU, dA are m x n column-major (ld m) at b*m*n; V, dV are
  // n x n (ld n) at b*n*n; s, ds at b*n.
  using V4 = typename Mfma<T>::V4;
  extern __shared__ double smem_[];
  const int ld = n16 + 1;
  T *const sJ = (T *)smem_;          // J, then N
  T *const sK = sJ + n16 * ld;       // K, then P
  T *const sR = sK + n16 * ld;       // 1/s_k (0 past n)
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
    // J = U^T dU (inner m) and K = V^T dV (inner n): tile (ti, tj) of either
    // is A(i, k) = X(k, ti*16 + i), B(k, j) = Y(k, tj*16 + j) with X, Y
    // column-major of column length len
    for (int task = w; task < 2 * nt * nt; task += BWD_WAVES) {
      const bool isK = task >= nt * nt;
      const int tt = isK ? task - nt * nt : task;
      const int ti = tt % nt, tj = tt / nt;
      const T *const X = isK ? Vb : Ub;
      const T *const Y = isK ? dVb : dUb;
      const int len = isK ? n : m;
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
      T *const dst = isK ? sK : sJ;
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
    // P = N V^T in place of K: A(i, k) = N(ti*16 + i, k) from LDS, B(k, c) =
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
      // through the wave's tile: lane l then stores rows 4(l%4)..+3 of
      // column l/4, so each column's 16 rows go out as one contiguous run
#pragma unroll
      for (int r = 0; r < 4; ++r)
        sT[li * TILE_LD + Mfma<T>::row(lk, r)] = acc[r];
      wave_sync();
      const int cc = lane >> 2, rr = 4 * (lane & 3);
      const int col = cj * 16 + cc;
      if (col < n) {
        T *const out = dA + (long long)b * m * n + (long long)col * m + ri * 16 + rr;
#pragma unroll
        for (int q = 0; q < 4; ++q)
          if (ri * 16 + rr + q < m)
            out[q] = sT[cc * TILE_LD + rr + q];
      }
      wave_sync();
    }
    __syncthreads();
  }
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
  const size_t lds = bwd_lds_bytes<T>(n16);
  if (lds > 160 * 1024)
    return set_error(-2, "svd backward: n = %d needs %zu bytes of LDS (at most 160 KiB; n <= 96 for FP64)", n, lds);
  int dev = 0, cus = 0;
  VJ_HIP(hipGetDevice(&dev));
  VJ_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
  const int per_cu = (int)std::max<size_t>(1, std::min<size_t>(8, (160 * 1024) / lds));
  const int grid = nb < per_cu * cus ? nb : per_cu * cus;
  if (lds > 64 * 1024)
    VJ_HIP(hipFuncSetAttribute((const void *)k_svd_bwd<T>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
  hipLaunchKernelGGL(k_svd_bwd<T>, dim3(grid), dim3(64 * BWD_WAVES), lds, st, nb, m, n, n16, U, s, V, dU, ds, dV, dA);
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

} // extern "C"
