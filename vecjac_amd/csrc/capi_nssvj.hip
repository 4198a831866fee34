// LAPACK xGESVJ-style drivers over the GPU Jacobi SVD: dnssvj_, snssvj_,
// znssvj_, cnssvj_ with the prototypes of the reference's dnssvj.f:335,
// snssvj.f:321, znssvj.f:349 and cnssvj.f:349 (LAPACK's xGESVJ with an
// NSWEEP argument), as the reference's tdgesvj.c:7 / tzgesvj.c:8 call them.
//
// What is kept from the LAPACK routine is its contract: the argument checks
// and INFO codes in its order (dnssvj.f:402-431, znssvj.f:416-446, with the
// complex workspace query), the convergence threshold TOL = CTOL * EPS with
// CTOL = sqrt(M) (JOBU = 'U' or a V requested), M (singular values only) or
// the caller's WORK(1) (JOBU = 'C'), the quick returns for a void, a zero and
// a one-column matrix (dnssvj.f:433-435, 595-621), the outputs -- singular
// values in non-increasing order as SCALE * SVA(1:N), the first N2 columns of
// A normalised into left singular vectors, V orthogonal or the rotations
// applied to an MV x N V (JOBV = 'A') -- and the six statistics in WORK(1:6)
// (dnssvj.f:1518-1600).  The rotations themselves are the library's parallel
// one-sided Jacobi (the same kernels as dvjsvd_, strategy table pjs id 2), not
// LAPACK's row-cyclic de Rijk pivoting, so singular vectors of clustered
// singular values and the sweep counts differ from LAPACK's, within the
// accuracy both guarantee.
//
// Shapes the GPU driver does not take are padded with zeros: an odd N gets
// a zero column N+1 and, when M = N, a zero row (neither changes the
// singular values nor meets a rotation: a zero column is orthogonal to every
// column and never out of order against the column before it).  JOBA = 'L' /
// 'U' reads the strict upper / lower triangle as zero, which is what the
// LAPACK routine assumes of it.
#include "svd_common.hpp"
#include "vecjac.h"

#include <algorithm>
#include <cmath>
#include <limits.h>
#include <limits>
#include <memory>
#include <numeric>
#include <vector>

using namespace vj;

namespace {

// LSAME: case-insensitive comparison of the first character
inline bool lsame(const char *c, char u) { return c && ((*c & ~0x20) == u); }

enum { TRI_G = 0, TRI_L = 1, TRI_U = 2 };

// Host layout (rows x cols, leading dimension ld, interleaved (re, im) when
// CPLX) -> split planes of rows_p x cols_p, zero outside the host block and
// in the triangle JOBA leaves out.  One workgroup per column.
template <typename T, bool CPLX>
__global__ __launch_bounds__(256) void k_nssvj_load(const T *__restrict__ src, const long long rows, const long long cols,
                                                    const long long ld, const long long rows_p, const int tri,
                                                    T *__restrict__ dr, T *__restrict__ di, const long long ldd)
{
  const long long j = blockIdx.x;
  for (long long i = threadIdx.x; i < rows_p; i += 256) {
    const bool in = (i < rows) && (j < cols) && !((tri == TRI_L && i < j) || (tri == TRI_U && i > j));
    T re = (T)0, im = (T)0;
    if (in) {
      if constexpr (CPLX) {
        re = src[2 * (j * ld + i)];
        im = src[2 * (j * ld + i) + 1];
      }
      else
        re = src[j * ld + i];
    }
    dr[j * ldd + i] = re;
    if constexpr (CPLX)
      di[j * ldd + i] = im;
  }
}

// Output column j := working column perm[j], either normalised by its norm
// fS * 2^eS (j < nnorm: the same operation as the epilogue of dvjsvd_) or
// multiplied by 2^shift; written compact (rows x gridDim.x, interleaved when
// CPLX).
template <typename T, bool CPLX>
__global__ __launch_bounds__(256) void k_nssvj_store(const T *__restrict__ gr, const T *__restrict__ gi,
                                                     const long long ldg, const long long rows,
                                                     const int *__restrict__ perm, const T *__restrict__ eS,
                                                     const T *__restrict__ fS, const int nnorm, const int shift,
                                                     T *__restrict__ dst)
{
  const long long j = blockIdx.x;
  const long long p = perm[j];
  const bool nrm = (j < nnorm);
  const T f_ = nrm ? fS[p] : (T)1, s_ = nrm ? -eS[p] : (T)0;
  for (long long i = threadIdx.x; i < rows; i += 256) {
    T re = gr[p * ldg + i];
    re = nrm ? scalef(re / f_, s_) : ldexp_(re, shift);
    if constexpr (CPLX) {
      T im = gi[p * ldg + i];
      im = nrm ? scalef(im / f_, s_) : ldexp_(im, shift);
      dst[2 * (j * rows + i)] = re;
      dst[2 * (j * rows + i) + 1] = im;
    }
    else
      dst[j * rows + i] = re;
  }
}

// max or sum over a workgroup of 256 lanes; every lane gets the result
__device__ __forceinline__ double block_reduce(double v, bool is_max, double *red)
{
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) {
    const double w = __shfl_xor(v, o, 64);
    v = is_max ? fmax(v, w) : (v + w);
  }
  __syncthreads();
  if ((threadIdx.x & 63) == 0)
    red[threadIdx.x >> 6] = v;
  __syncthreads();
  v = red[0];
  for (int w = 1; w < 4; ++w)
    v = is_max ? fmax(v, red[w]) : (v + red[w]);
  return v;
}

// Column norms as (e, f) pairs, norm = f * 2^e with f in [1, 2) (zero:
// e = -inf, f = 1), for a run that stopped at NSWEEP: the norms the sweeps
// leave behind are stale for the columns rotated in the last step.  Scaled
// sum of squares in double: max |g| = 2^kx * [1, 2), exact power-of-two
// scaling, one square root.
template <typename T, bool CPLX>
__global__ __launch_bounds__(256) void k_nssvj_norms(const T *__restrict__ gr, const T *__restrict__ gi,
                                                     const long long ldg, const long long rows, T *__restrict__ eS,
                                                     T *__restrict__ fS)
{
  __shared__ double red[4];
  const long long j = blockIdx.x;
  double mx = 0.0;
  for (long long i = threadIdx.x; i < rows; i += 256) {
    mx = fmax(mx, fabs((double)gr[j * ldg + i]));
    if constexpr (CPLX)
      mx = fmax(mx, fabs((double)gi[j * ldg + i]));
  }
  mx = block_reduce(mx, true, red);
  if (mx == 0.0) {
    if (threadIdx.x == 0) {
      eS[j] = -inf_<T>();
      fS[j] = (T)1;
    }
    return;
  }
  const int kx = ilogb(mx);
  double ss = 0.0;
  for (long long i = threadIdx.x; i < rows; i += 256) {
    const double x = ldexp((double)gr[j * ldg + i], -kx);
    ss = fma(x, x, ss);
    if constexpr (CPLX) {
      const double y = ldexp((double)gi[j * ldg + i], -kx);
      ss = fma(y, y, ss);
    }
  }
  ss = block_reduce(ss, false, red);
  if (threadIdx.x == 0) {
    int e = 0;
    const double f = 2.0 * frexp(sqrt(ss), &e);
    T ft = (T)f;
    int et = kx + e - 1;
    if (ft >= (T)2) { // rounded up to T
      ft = (T)1;
      ++et;
    }
    eS[j] = (T)et;
    fS[j] = ft;
  }
}

template <typename T> struct Lapack;
template <> struct Lapack<double> {
  static constexpr int EMAX = 1023, EMIN = -1022; // exponents of the normal range
};
template <> struct Lapack<float> {
  static constexpr int EMAX = 127, EMIN = -126;
};

struct Job {
  bool lsvec, uctol, rsvec, applv;
  int tri;
};

// One singular value sigma = f * 2^e (f in [1, 2); e = -inf for zero).
template <typename T> struct EF {
  double e;
  T f;
};
template <typename T> inline bool ef_greater(const EF<T> &a, const EF<T> &b)
{
  return (a.e > b.e) || ((a.e == b.e) && (a.f > b.f));
}

// Everything after the argument checks.  A and V are host arrays (T or
// interleaved complex), stat is WORK (real) / RWORK (complex); returns INFO.
template <typename T, bool CPLX>
vj_fint run_nssvj(const Job &jb, const vj_fint M, const vj_fint N, T *A, const vj_fint lda, T *sva, const vj_fint MV,
                  T *V, const vj_fint ldv, T *stat, const vj_fint nsweep)
{
  constexpr int W = CPLX ? 2 : 1; // T words per element
  const int L_MG = CPLX ? 92 : 80, L_MV = CPLX ? 104 : 92;
  if (M == 0 || N == 0)
    return 0;
  const T eps = Lim<T>::EPS / (T)2; // xLAMCH('Epsilon')
  if ((T)M * eps >= (T)1)
    return -4;
  if (nsweep < 1)
    return set_error(CPLX ? -16 : -14, "NSWEEP = %lld: at least one sweep is required", (long long)nsweep);
  if (M > 0x3fffffffLL || N > 0x3fffffffLL)
    return set_error(-4, "M = %lld, N = %lld exceed the 32-bit index range of the kernels", (long long)M, (long long)N);
  const T ctol = jb.uctol ? stat[0] : ((jb.lsvec || jb.rsvec || jb.applv) ? std::sqrt((T)M) : (T)M);
  const long long Np = N + (N & 1), Mp = std::max<long long>(M, Np);
  // JOBV = 'A' with an empty or all-zero V: the product stays zero (V is
  // not written); the SVD runs with an internal V
  bool vpre = jb.applv && MV > 0;
  if (vpre) {
    bool nz = false;
    for (vj_fint j = 0; j < N && !nz; ++j)
      for (vj_fint i = 0; i < W * MV; ++i)
        if (V[(size_t)j * ldv * W + i] != (T)0) {
          nz = true;
          break;
        }
    vpre = nz;
  }
  const long long nv = vpre ? (long long)MV : Np;

  hipStream_t s = lib_stream();
  Staging sg(s);
  PlainArgs<T> a{};
  a.m = Mp;
  a.n = Np;
  a.ldG = Mp;
  a.ldV = nv;
  a.nv = nv;
  a.vpreset = vpre ? 1 : 0;
  a.tol = ctol * eps;
  a.raw = 1;
  a.stats = 1;
  a.swp = (unsigned)std::min<vj_fint>(nsweep, (vj_fint)INT_MAX);
  unsigned stp = 0u;
  unsigned *tab = pjs(2l, (unsigned)Np, &stp);
  if (!tab)
    return set_error(VJ_EHIP, "strategy table allocation failed");
  std::unique_ptr<unsigned, void (*)(unsigned *)> tab_guard(tab, pjs_free);
  a.stp = stp;
  T *stage = nullptr, *vstage = nullptr;
  unsigned *djs = nullptr;
  int *dperm = nullptr;
  if (int rc = sg.alloc(&a.Gr, (size_t)Mp * Np)) return rc;
  if (int rc = sg.alloc(&a.Vr, (size_t)nv * Np)) return rc;
  if (CPLX) {
    if (int rc = sg.alloc(&a.Gi, (size_t)Mp * Np)) return rc;
    if (int rc = sg.alloc(&a.Vi, (size_t)nv * Np)) return rc;
  }
  if (int rc = sg.alloc(&a.eS, (size_t)Np)) return rc;
  if (int rc = sg.alloc(&a.fS, (size_t)Np)) return rc;
  if (int rc = sg.alloc(&djs, (size_t)stp * Np)) return rc;
  if (int rc = sg.alloc(&dperm, (size_t)N)) return rc;
  if (int rc = sg.alloc(&stage, (size_t)W * M * N)) return rc;
  if (vpre)
    if (int rc = sg.alloc(&vstage, (size_t)W * MV * N)) return rc;
  a.js = djs;
  const size_t eb = W * sizeof(T);
  VJ_HIP(hipMemcpyAsync(djs, tab, (size_t)stp * Np * sizeof(unsigned), hipMemcpyHostToDevice, s));
  VJ_HIP(hipMemcpy2DAsync(stage, (size_t)M * eb, A, (size_t)lda * eb, (size_t)M * eb, (size_t)N, hipMemcpyHostToDevice, s));
  hipLaunchKernelGGL((k_nssvj_load<T, CPLX>), dim3((unsigned)Np), dim3(256), 0, s, stage, (long long)M, (long long)N,
                     (long long)M, Mp, jb.tri, a.Gr, a.Gi, Mp);
  if (vpre) {
    VJ_HIP(hipMemcpy2DAsync(vstage, (size_t)MV * eb, V, (size_t)ldv * eb, (size_t)MV * eb, (size_t)N,
                            hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL((k_nssvj_load<T, CPLX>), dim3((unsigned)Np), dim3(256), 0, s, vstage, (long long)MV,
                       (long long)N, (long long)MV, nv, (int)TRI_G, a.Vr, a.Vi, nv);
  }
  if (int rc = check_launch())
    return rc;
  PlainRun run{};
  const int rc = plain_svd<T, CPLX>(a, &run, s);
  if (rc == -L_MG)
    return set_error(-6, "A has an infinite or NaN entry");
  if (vpre && (rc == -L_MV))
    return set_error(-10, "V has an infinite or NaN entry");
  if (rc < 0)
    return rc;
  // shifts applied to the working G and V (a rescaling decided at the last
  // step of a run that stopped at NSWEEP is counted but not yet applied)
  const int sT = run.sT - run.sG_pending, sV = run.sV - run.sV_pending;
  const bool conv = run.sweeps < a.swp;

  // zero matrix (dnssvj.f:599-608): U := I if requested, SVA = 0, V = I
  if (run.zero) {
    for (vj_fint j = 0; j < N; ++j)
      sva[j] = (T)0;
    if (jb.lsvec)
      for (vj_fint j = 0; j < N; ++j)
        for (vj_fint i = 0; i < W * M; ++i)
          A[(size_t)j * lda * W + i] = (i == W * j) ? (T)1 : (T)0;
    if (jb.rsvec)
      for (vj_fint j = 0; j < N; ++j)
        for (vj_fint i = 0; i < W * N; ++i)
          V[(size_t)j * ldv * W + i] = (i == W * j) ? (T)1 : (T)0;
    stat[0] = (T)1;
    for (int k = 1; k < 6; ++k)
      stat[k] = (T)0;
    return 0;
  }

  // singular values sigma_j = fS_j 2^(eS_j - sT), sorted non-increasingly;
  // a converged run's last sweep rotated nothing and left every norm fresh
  if (!conv) {
    hipLaunchKernelGGL((k_nssvj_norms<T, CPLX>), dim3((unsigned)Np), dim3(256), 0, s, a.Gr, a.Gi, Mp, Mp, a.eS, a.fS);
    if (int rc2 = check_launch())
      return rc2;
  }
  std::vector<T> heS((size_t)Np), hfS((size_t)Np);
  VJ_HIP(hipMemcpyAsync(heS.data(), a.eS, (size_t)Np * sizeof(T), hipMemcpyDeviceToHost, s));
  VJ_HIP(hipMemcpyAsync(hfS.data(), a.fS, (size_t)Np * sizeof(T), hipMemcpyDeviceToHost, s));
  VJ_HIP(hipStreamSynchronize(s));
  std::vector<EF<T>> sg_((size_t)N);
  for (vj_fint j = 0; j < N; ++j)
    sg_[j] = EF<T>{(double)heS[j] - (double)sT, hfS[j]};
  std::vector<int> perm((size_t)N);
  std::iota(perm.begin(), perm.end(), 0);
  std::stable_sort(perm.begin(), perm.end(), [&](int x, int y) { return ef_greater(sg_[x], sg_[y]); });
  // N4: nonzero values, N2: values above the underflow threshold
  vj_fint n4 = 0, n2 = 0;
  double emax = -INFINITY, emin = INFINITY;
  for (vj_fint j = 0; j < N; ++j) {
    const EF<T> &v = sg_[j];
    if (v.e == -INFINITY)
      continue;
    ++n4;
    if ((v.e > (double)Lapack<T>::EMIN) || ((v.e == (double)Lapack<T>::EMIN) && (v.f > (T)1)))
      ++n2;
    emax = std::max(emax, v.e);
    emin = std::min(emin, v.e);
  }
  // SCALE = 2^k: 1 when every nonzero value is a normal number, else the
  // shift that keeps the largest finite and lifts the smallest as far as
  // that allows
  int k = 0;
  if (n4) {
    if (emax > Lapack<T>::EMAX)
      k = (int)(emax - Lapack<T>::EMAX);
    else if (emin < Lapack<T>::EMIN)
      k = std::max((int)(emin - Lapack<T>::EMIN), (int)(emax - Lapack<T>::EMAX));
    k = std::max(k, Lapack<T>::EMIN);
  }
  for (vj_fint j = 0; j < N; ++j) {
    const EF<T> &v = sg_[perm[j]];
    sva[j] = (v.e == -INFINITY) ? (T)0 : std::ldexp(v.f, (int)(v.e - k));
  }
  VJ_HIP(hipMemcpyAsync(dperm, perm.data(), (size_t)N * sizeof(int), hipMemcpyHostToDevice, s));

  // A: normalised left singular vectors for the first N2 columns (JOBU 'U',
  // 'F' or 'C'), the columns of U * SVA otherwise; one-column matrices are
  // normalised for 'U' / 'F' only (dnssvj.f:612-614)
  const bool one = (N == 1);
  const bool write_a = one ? jb.lsvec : true;
  if (write_a) {
    const int nnorm = (jb.lsvec || jb.uctol) ? (int)n2 : 0;
    hipLaunchKernelGGL((k_nssvj_store<T, CPLX>), dim3((unsigned)N), dim3(256), 0, s, a.Gr, a.Gi, Mp, (long long)M,
                       dperm, a.eS, a.fS, nnorm, -(sT + k), stage);
    if (int rc2 = check_launch())
      return rc2;
    VJ_HIP(hipMemcpy2DAsync(A, (size_t)lda * eb, stage, (size_t)M * eb, (size_t)M * eb, (size_t)N,
                            hipMemcpyDeviceToHost, s));
  }
  // V: the accumulated rotations (JOBV 'V'), or applied to the MV x N V
  const bool write_v = jb.rsvec || vpre;
  if (write_v) {
    const long long vrows = vpre ? (long long)MV : (long long)N;
    // stage (M x N >= N x N elements) is free again once the copy of A
    // before this launch on the same stream has ended
    T *vout = vpre ? vstage : stage;
    hipLaunchKernelGGL((k_nssvj_store<T, CPLX>), dim3((unsigned)N), dim3(256), 0, s, a.Vr, a.Vi, nv, vrows, dperm,
                       a.eS, a.fS, 0, -sV, vout);
    if (int rc2 = check_launch())
      return rc2;
    VJ_HIP(hipMemcpy2DAsync(V, (size_t)ldv * eb, vout, (size_t)vrows * eb, (size_t)vrows * eb, (size_t)N,
                            hipMemcpyDeviceToHost, s));
  }
  VJ_HIP(hipStreamSynchronize(s));

  stat[0] = std::ldexp((T)1, k);
  if (one) {
    // dnssvj.f:612-621
    stat[1] = (sva[0] * stat[0] >= std::numeric_limits<T>::min()) ? (T)1 : (T)0;
    for (int q = 2; q < 6; ++q)
      stat[q] = (T)0;
    return 0;
  }
  stat[1] = (T)n4;
  stat[2] = (T)n2;
  // the sweep index at exit: the converged sweep, or NSWEEP + 1 (the DO
  // variable after the loop ran to its end)
  stat[3] = conv ? (T)(run.sweeps + 1u) : (T)((double)nsweep + 1.0);
  stat[4] = (T)run.last_cos;
  stat[5] = (T)run.last_sin;
  return conv ? 0 : nsweep - 1;
}

Job make_job(const char *joba, const char *jobu, const char *jobv, bool cplx)
{
  Job j{};
  j.lsvec = lsame(jobu, 'U') || (cplx && lsame(jobu, 'F'));
  j.uctol = lsame(jobu, 'C');
  j.rsvec = lsame(jobv, 'V') || (cplx && lsame(jobv, 'J'));
  j.applv = lsame(jobv, 'A');
  j.tri = lsame(joba, 'L') ? TRI_L : (lsame(joba, 'U') ? TRI_U : TRI_G);
  return j;
}

// dnssvj.f:402-431 / snssvj.f:395-416
template <typename T>
vj_fint check_real(const char *joba, const char *jobu, const char *jobv, vj_fint m, vj_fint n, vj_fint lda,
                   vj_fint mv, vj_fint ldv, const T *work, vj_fint lwork, const Job &jb)
{
  if (!(jb.tri != TRI_G || lsame(joba, 'G'))) return -1;
  if (!(jb.lsvec || jb.uctol || lsame(jobu, 'N'))) return -2;
  if (!(jb.rsvec || jb.applv || lsame(jobv, 'N'))) return -3;
  if (m < 0) return -4;
  if (n < 0 || n > m) return -5;
  if (lda < m) return -7;
  if (mv < 0) return -9;
  if ((jb.rsvec && ldv < n) || (jb.applv && ldv < mv)) return -11;
  if (jb.uctol && work[0] <= (T)1) return -12;
  if (lwork < std::max<vj_fint>(m + n, 6)) return -13;
  return 0;
}

// znssvj.f:416-446 / cnssvj.f:416-446; 1 = workspace query
template <typename T>
vj_fint check_cplx(const char *joba, const char *jobu, const char *jobv, vj_fint m, vj_fint n, vj_fint lda,
                   vj_fint mv, vj_fint ldv, vj_fint lwork, const T *rwork, vj_fint lrwork, const Job &jb)
{
  const bool lquery = (lwork == -1) || (lrwork == -1);
  if (!(jb.tri != TRI_G || lsame(joba, 'G'))) return -1;
  if (!(jb.lsvec || jb.uctol || lsame(jobu, 'N'))) return -2;
  if (!(jb.rsvec || jb.applv || lsame(jobv, 'N'))) return -3;
  if (m < 0) return -4;
  if (n < 0 || n > m) return -5;
  if (lda < m) return -7;
  if (mv < 0) return -9;
  if ((jb.rsvec && ldv < n) || (jb.applv && ldv < mv)) return -11;
  if (jb.uctol && rwork[0] <= (T)1) return -12;
  if (lwork < m + n && !lquery) return -13;
  if (lrwork < std::max<vj_fint>(n, 6) && !lquery) return -15;
  return lquery ? 1 : 0;
}

template <typename T>
void real_entry(const char *joba, const char *jobu, const char *jobv, const vj_fint *m, const vj_fint *n, T *A,
                const vj_fint *lda, T *sva, const vj_fint *mv, T *V, const vj_fint *ldv, T *work, const vj_fint *lwork,
                const vj_fint *nsweep, vj_fint *info)
{
  const Job jb = make_job(joba, jobu, jobv, false);
  *info = check_real<T>(joba, jobu, jobv, *m, *n, *lda, *mv, *ldv, work, *lwork, jb);
  if (*info) {
    (void)set_error((int)*info, "xNSSVJ: parameter number %lld had an illegal value", (long long)-*info);
    return;
  }
  *info = run_nssvj<T, false>(jb, *m, *n, A, *lda, sva, *mv, V, *ldv, work, *nsweep);
}

template <typename T>
void cplx_entry(const char *joba, const char *jobu, const char *jobv, const vj_fint *m, const vj_fint *n, T *A,
                const vj_fint *lda, T *sva, const vj_fint *mv, T *V, const vj_fint *ldv, T *cwork,
                const vj_fint *lwork, T *rwork, const vj_fint *lrwork, const vj_fint *nsweep, vj_fint *info)
{
  const Job jb = make_job(joba, jobu, jobv, true);
  const vj_fint c = check_cplx<T>(joba, jobu, jobv, *m, *n, *lda, *mv, *ldv, *lwork, rwork, *lrwork, jb);
  if (c < 0) {
    *info = c;
    (void)set_error((int)c, "xNSSVJ: parameter number %lld had an illegal value", (long long)-c);
    return;
  }
  if (c == 1) { // workspace query (znssvj.f:451-455)
    *info = 0;
    cwork[0] = (T)(*m + *n);
    cwork[1] = (T)0;
    rwork[0] = (T)std::max<vj_fint>(*n, 6);
    return;
  }
  *info = run_nssvj<T, true>(jb, *m, *n, A, *lda, sva, *mv, V, *ldv, rwork, *nsweep);
}

} // namespace

extern "C" {

void dnssvj_(const char *joba, const char *jobu, const char *jobv, const vj_fint *m, const vj_fint *n, double *A,
             const vj_fint *lda, double *sva, const vj_fint *mv, double *V, const vj_fint *ldv, double *work,
             const vj_fint *lwork, const vj_fint *nsweep, vj_fint *info)
{
  real_entry<double>(joba, jobu, jobv, m, n, A, lda, sva, mv, V, ldv, work, lwork, nsweep, info);
}

void snssvj_(const char *joba, const char *jobu, const char *jobv, const vj_fint *m, const vj_fint *n, float *A,
             const vj_fint *lda, float *sva, const vj_fint *mv, float *V, const vj_fint *ldv, float *work,
             const vj_fint *lwork, const vj_fint *nsweep, vj_fint *info)
{
  real_entry<float>(joba, jobu, jobv, m, n, A, lda, sva, mv, V, ldv, work, lwork, nsweep, info);
}

void znssvj_(const char *joba, const char *jobu, const char *jobv, const vj_fint *m, const vj_fint *n, vj_zdouble *A,
             const vj_fint *lda, double *sva, const vj_fint *mv, vj_zdouble *V, const vj_fint *ldv, vj_zdouble *cwork,
             const vj_fint *lwork, double *rwork, const vj_fint *lrwork, const vj_fint *nsweep, vj_fint *info)
{
  cplx_entry<double>(joba, jobu, jobv, m, n, A, lda, sva, mv, V, ldv, cwork, lwork, rwork, lrwork, nsweep, info);
}

void cnssvj_(const char *joba, const char *jobu, const char *jobv, const vj_fint *m, const vj_fint *n, vj_cfloat *A,
             const vj_fint *lda, float *sva, const vj_fint *mv, vj_cfloat *V, const vj_fint *ldv, vj_cfloat *cwork,
             const vj_fint *lwork, float *rwork, const vj_fint *lrwork, const vj_fint *nsweep, vj_fint *info)
{
  cplx_entry<float>(joba, jobu, jobv, m, n, A, lda, sva, mv, V, ldv, cwork, lwork, rwork, lrwork, nsweep, info);
}

} // extern "C"
