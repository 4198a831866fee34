// Adjoint (backward) of the thin real SVD A = U diag(s) V^T for a batch of
// matrices (no reference counterpart).  With F_ij = 1/(s_j^2 - s_i^2) for
// i != j (0 on the diagonal), J = U^T dU and K = V^T dV:
//   dA = U [diag(ds) + (F o (J - J^T)) S + S (F o (K - K^T))] V^T
//        + (I - U U^T) dU S^-1 V^T
// which is evaluated as dA = (U N + dU S^-1) V^T with
//   N = diag(ds) + (F o (J - J^T)) S + S (F o (K - K^T)) - J S^-1.
// (V is square, so the (I - V V^T) term vanishes.)  One workgroup per
// matrix; U, dU, V and the n x n intermediates are staged in LDS (64 KiB
// per workgroup at most).
#include "svd_common.hpp"
#include "vecjac.h"

namespace vj {

template <typename T>
__global__ __launch_bounds__(256) void k_svd_bwd(const int nb, const int m, const int n, const T *__restrict__ U,
                                                 const T *__restrict__ s, const T *__restrict__ V,
                                                 const T *__restrict__ dU, const T *__restrict__ ds,
                                                 const T *__restrict__ dV, T *__restrict__ dA)
{
  // per matrix b: U, dU, dA are m x n column-major (ld m) at b*m*n; V, dV are
  // n x n (ld n) at b*n*n; s, ds at b*n.
  extern __shared__ double smem[];
  T *const sU = (T *)smem;            // m*n
  T *const sdU = sU + m * n;          // m*n
  T *const sV = sdU + m * n;          // n*n
  T *const sJ = sV + n * n;           // n*n
  T *const sN = sJ + n * n;           // n*n
  T *const ss = sN + n * n;           // n
  T *const sds = ss + n;              // n
  for (int b = blockIdx.x; b < nb; b += gridDim.x) {
    const long long oU = (long long)b * m * n, oV = (long long)b * n * n, oS = (long long)b * n;
    // dV is staged in sJ until K has been formed
    for (int t = threadIdx.x; t < m * n; t += blockDim.x) {
      sU[t] = U[oU + t];
      sdU[t] = dU[oU + t];
    }
    for (int t = threadIdx.x; t < n * n; t += blockDim.x) {
      sV[t] = V[oV + t];
      sJ[t] = dV[oV + t];
    }
    for (int t = threadIdx.x; t < n; t += blockDim.x) {
      ss[t] = s[oS + t];
      sds[t] = ds[oS + t];
    }
    __syncthreads();
    // K = V^T dV into sN.  Lane (i, j) runs its sum from row i on (r' =
    // (r + i) mod len), so that the lanes of a wave, which differ in i, read
    // different LDS banks instead of one column stride apart.
    for (int t = threadIdx.x; t < n * n; t += blockDim.x) {
      const int i = t % n, j = t / n; // element (i, j), column-major
      T acck = (T)0;
      int r = i % n;
      for (int k = 0; k < n; ++k) {
        acck = fma_(sV[i * n + r], sJ[j * n + r], acck);
        r = (r + 1 == n) ? 0 : (r + 1);
      }
      sN[t] = acck;
    }
    __syncthreads();
    // J = U^T dU into sJ (dV is no longer needed)
    T jv[16];
    int cj = 0;
    for (int t = threadIdx.x; t < n * n; t += blockDim.x) {
      const int i = t % n, j = t / n;
      T acc = (T)0;
      int r = i % m;
      for (int k = 0; k < m; ++k) {
        acc = fma_(sU[i * m + r], sdU[j * m + r], acc);
        r = (r + 1 == m) ? 0 : (r + 1);
      }
      jv[cj++] = acc;
    }
    cj = 0;
    for (int t = threadIdx.x; t < n * n; t += blockDim.x)
      sJ[t] = jv[cj++];
    __syncthreads();
    // N_ij
    T nv[16];
    int cnt = 0;
    for (int t = threadIdx.x; t < n * n; t += blockDim.x) {
      const int i = t % n, j = t / n;
      T v;
      if (i == j)
        v = sds[i] - sJ[t] / ss[j];
      else {
        const T si = ss[i], sj = ss[j];
        const T F = (T)1 / ((sj - si) * (sj + si));
        // column-major: (i, j) at t = i + j n, (j, i) at j + i n
        const T aJ = sJ[t] - sJ[i * n + j];
        const T aK = sN[t] - sN[i * n + j];
        v = F * (aJ * sj + si * aK) - sJ[t] / sj;
      }
      nv[cnt++] = v;
    }
    __syncthreads();
    cnt = 0;
    for (int t = threadIdx.x; t < n * n; t += blockDim.x)
      sN[t] = nv[cnt++];
    __syncthreads();
    // X = U N + dU S^-1, in place of dU (element t reads only dU[t])
    for (int t = threadIdx.x; t < m * n; t += blockDim.x) {
      const int r = t % m, j = t / m;
      T acc = sdU[t] / ss[j];
      for (int k = 0; k < n; ++k)
        acc = fma_(sU[k * m + r], sN[j * n + k], acc);
      sdU[t] = acc;
    }
    __syncthreads();
    // dA = X V^T: dA(r, c) = sum_j X(r, j) V(c, j)
    for (int t = threadIdx.x; t < m * n; t += blockDim.x) {
      const int r = t % m, c = t / m;
      T acc = (T)0;
      for (int k = 0; k < n; ++k)
        acc = fma_(sdU[k * m + r], sV[k * n + c], acc);
      dA[oU + t] = acc;
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
  const size_t lds = (size_t)(2 * m * n + 3 * n * n + 2 * n) * sizeof(T);
  if (lds > 64 * 1024 || n * n > 16 * 256)
    return set_error(-2, "svd backward: m x n = %d x %d does not fit in LDS", m, n);
  int dev = 0, cus = 0;
  VJ_HIP(hipGetDevice(&dev));
  VJ_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
  const int grid = nb < 8 * cus ? nb : 8 * cus;
  hipLaunchKernelGGL(k_svd_bwd<T>, dim3(grid), dim3(256), lds, st, nb, m, n, U, s, V, dU, ds, dV, dA);
  return check_launch();
}

template int svd_bwd<double>(int, int, int, const double *, const double *, const double *, const double *, const double *,
                             const double *, double *, hipStream_t);
template int svd_bwd<float>(int, int, int, const float *, const float *, const float *, const float *, const float *,
                            const float *, float *, hipStream_t);

} // namespace vj

extern "C" {

// dA for a batch of nb thin SVDs of m x n matrices; see the header.
vj_fint vjh_dsvd_bwd(vj_fint nb, vj_fint m, vj_fint n, const double *U, const double *s, const double *V,
                     const double *dU, const double *ds, const double *dV, double *dA, void *stream)
{
  return vj::svd_bwd<double>((int)nb, (int)m, (int)n, U, s, V, dU, ds, dV, dA, (hipStream_t)stream);
}
vj_fint vjh_ssvd_bwd(vj_fint nb, vj_fint m, vj_fint n, const float *U, const float *s, const float *V,
                     const float *dU, const float *ds, const float *dV, float *dA, void *stream)
{
  return vj::svd_bwd<float>((int)nb, (int)m, (int)n, U, s, V, dU, ds, dV, dA, (hipStream_t)stream);
}

} // extern "C"
