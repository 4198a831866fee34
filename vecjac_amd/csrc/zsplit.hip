// Interleaved <-> split complex layout (zsplit_ / zmerge_, zsplit.c:3 and
// zmerge.c:3): the complex SVD works on split real and imaginary planes.
//   zsplit_ / zmerge_       the reference's prototypes, host memory (a plain
//                           column loop: a layout change, no arithmetic);
//   vjh_zsplit / vjh_zmerge the same on device memory, one coalesced pass
//                           on a stream, so interleaved data already in HBM
//                           never takes a host round trip.
#include "capi_common.hpp"
#include "vecjac.h"

using namespace vj;

namespace {

// one thread per element: the interleaved element is one 16-byte load
__global__ __launch_bounds__(256) void k_zsplit(const long long m, const long long n, const double2 *__restrict__ A,
                                                const long long ldA, double *__restrict__ Ar, const long long ldAr,
                                                double *__restrict__ Ai, const long long ldAi)
{
  const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
  const long long j = blockIdx.y;
  if (i >= m || j >= n)
    return;
  const double2 a = A[j * ldA + i];
  Ar[j * ldAr + i] = a.x;
  Ai[j * ldAi + i] = a.y;
}

__global__ __launch_bounds__(256) void k_zmerge(const long long m, const long long n, const double *__restrict__ Ar,
                                                const long long ldAr, const double *__restrict__ Ai, const long long ldAi,
                                                double2 *__restrict__ A, const long long ldA)
{
  const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
  const long long j = blockIdx.y;
  if (i >= m || j >= n)
    return;
  A[j * ldA + i] = make_double2(Ar[j * ldAr + i], Ai[j * ldAi + i]);
}

// argument checks of the reference (zsplit.c:6-23, debug builds): m, the
// leading dimensions; alignment is not required here
int check(vjh_nat m, vjh_nat ldA, vjh_nat ldAr, vjh_nat ldAi)
{
  if (ldA < m)
    return -4;
  if (ldAr < m)
    return -6;
  if (ldAi < m)
    return -8;
  return 0;
}

int launch_dims(vjh_nat m, vjh_nat n, dim3 &g)
{
  if (n > 65535u)
    return set_error(-2, "n = %llu exceeds the grid limit of the split kernels", (unsigned long long)n);
  g = dim3((unsigned)((m + 255u) / 256u), (unsigned)n);
  return 0;
}

} // namespace

extern "C" {

vj_fint zsplit_(const vj_fnat *m, const vj_fnat *n, const vj_zdouble *A, const vj_fnat *ldA, double *Ar, const vj_fnat *ldAr,
                double *Ai, const vj_fnat *ldAi)
{
  if (int rc = check(*m, *ldA, *ldAr, *ldAi))
    return rc;
  const double *a = (const double *)A;
  for (vj_fnat j = 0; j < *n; ++j)
    for (vj_fnat i = 0; i < *m; ++i) {
      Ar[j * *ldAr + i] = a[2u * (j * *ldA + i)];
      Ai[j * *ldAi + i] = a[2u * (j * *ldA + i) + 1u];
    }
  return 0;
}

vj_fint zmerge_(const vj_fnat *m, const vj_fnat *n, const double *Ar, const vj_fnat *ldAr, const double *Ai,
                const vj_fnat *ldAi, vj_zdouble *A, const vj_fnat *ldA)
{
  if (*ldAr < *m)
    return -4;
  if (*ldAi < *m)
    return -6;
  if (*ldA < *m)
    return -8;
  double *a = (double *)A;
  for (vj_fnat j = 0; j < *n; ++j)
    for (vj_fnat i = 0; i < *m; ++i) {
      a[2u * (j * *ldA + i)] = Ar[j * *ldAr + i];
      a[2u * (j * *ldA + i) + 1u] = Ai[j * *ldAi + i];
    }
  return 0;
}

vjh_int vjh_zsplit(vjh_nat m, vjh_nat n, const double *A, vjh_nat ldA, double *Ar, vjh_nat ldAr, double *Ai, vjh_nat ldAi,
                   void *stream)
{
  if (int rc = check(m, ldA, ldAr, ldAi))
    return rc;
  if (!m || !n)
    return 0;
  dim3 g;
  if (int rc = launch_dims(m, n, g))
    return rc;
  hipLaunchKernelGGL(k_zsplit, g, dim3(256), 0, (hipStream_t)stream, (long long)m, (long long)n, (const double2 *)A,
                     (long long)ldA, Ar, (long long)ldAr, Ai, (long long)ldAi);
  return check_launch();
}

vjh_int vjh_zmerge(vjh_nat m, vjh_nat n, const double *Ar, vjh_nat ldAr, const double *Ai, vjh_nat ldAi, double *A, vjh_nat ldA,
                   void *stream)
{
  if (ldAr < m)
    return -4;
  if (ldAi < m)
    return -6;
  if (ldA < m)
    return -8;
  if (!m || !n)
    return 0;
  dim3 g;
  if (int rc = launch_dims(m, n, g))
    return rc;
  hipLaunchKernelGGL(k_zmerge, g, dim3(256), 0, (hipStream_t)stream, (long long)m, (long long)n, Ar, (long long)ldAr, Ai,
                     (long long)ldAi, (double2 *)A, (long long)ldA);
  return check_launch();
}

} // extern "C"
