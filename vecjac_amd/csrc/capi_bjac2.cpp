// C-ABI entry points of the batched 2x2 Hermitian EVD.
//   dbjac2_/zbjac2_/sbjac2_/cbjac2_: the reference's prototypes (dbjac2.h:6,
//     zbjac2.h:6, sbjac2.h:6, cbjac2.h:6), host arrays in and out;
//   vjh_?bjac2: the same computation on device arrays, enqueued on a stream.
// The sign of n selects sines (n<0) or tangents (n>0), as in the reference.
// |n| must be a multiple of the reference's lane count (8 for FP64, 16 for
// FP32) because p[] packs that many L1<L2 bits per word; otherwise -1 (the
// reference's own argument check, dbjac2.c:86-87).  Returns 0 on success.
#include "capi_common.hpp"
#include "vecjac.h"

using namespace vj;

namespace {

template <typename T>
int host_rbjac2(const vj_fint *n, const T *a11, const T *a22, const T *a21, T *c, T *at, T *l1, T *l2, unsigned *p)
{
  const long long nn = (*n < 0) ? -(long long)*n : (long long)*n;
  if (nn % Lim<T>::VL)
    return -1;
  if (!nn)
    return 0;
  hipStream_t s = lib_stream();
  Staging st(s);
  T *d[7];
  unsigned *dp = nullptr;
  for (int k = 0; k < 7; ++k)
    if (int rc = st.alloc(d + k, (size_t)nn))
      return rc;
  if (int rc = st.alloc(&dp, (size_t)(nn / Lim<T>::VL)))
    return rc;
  const size_t b = (size_t)nn * sizeof(T);
  VJ_HIP(hipMemcpyAsync(d[0], a11, b, hipMemcpyHostToDevice, s));
  VJ_HIP(hipMemcpyAsync(d[1], a22, b, hipMemcpyHostToDevice, s));
  VJ_HIP(hipMemcpyAsync(d[2], a21, b, hipMemcpyHostToDevice, s));
  if (int rc = launch_rbjac2<T>((long long)*n, d[0], d[1], d[2], d[3], d[4], d[5], d[6], dp, s))
    return rc;
  VJ_HIP(hipMemcpyAsync(c, d[3], b, hipMemcpyDeviceToHost, s));
  VJ_HIP(hipMemcpyAsync(at, d[4], b, hipMemcpyDeviceToHost, s));
  VJ_HIP(hipMemcpyAsync(l1, d[5], b, hipMemcpyDeviceToHost, s));
  VJ_HIP(hipMemcpyAsync(l2, d[6], b, hipMemcpyDeviceToHost, s));
  VJ_HIP(hipMemcpyAsync(p, dp, (size_t)(nn / Lim<T>::VL) * sizeof(unsigned), hipMemcpyDeviceToHost, s));
  VJ_HIP(hipStreamSynchronize(s));
  return 0;
}

template <typename T>
int host_cbjac2(const vj_fint *n, const T *a11, const T *a22, const T *a21r, const T *a21i, T *c, T *cat, T *sat, T *l1, T *l2, unsigned *p)
{
  const long long nn = (*n < 0) ? -(long long)*n : (long long)*n;
  if (nn % Lim<T>::VL)
    return -1;
  if (!nn)
    return 0;
  hipStream_t s = lib_stream();
  Staging st(s);
  T *d[9];
  unsigned *dp = nullptr;
  for (int k = 0; k < 9; ++k)
    if (int rc = st.alloc(d + k, (size_t)nn))
      return rc;
  if (int rc = st.alloc(&dp, (size_t)(nn / Lim<T>::VL)))
    return rc;
  const size_t b = (size_t)nn * sizeof(T);
  VJ_HIP(hipMemcpyAsync(d[0], a11, b, hipMemcpyHostToDevice, s));
  VJ_HIP(hipMemcpyAsync(d[1], a22, b, hipMemcpyHostToDevice, s));
  VJ_HIP(hipMemcpyAsync(d[2], a21r, b, hipMemcpyHostToDevice, s));
  VJ_HIP(hipMemcpyAsync(d[3], a21i, b, hipMemcpyHostToDevice, s));
  if (int rc = launch_cbjac2<T>((long long)*n, d[0], d[1], d[2], d[3], d[4], d[5], d[6], d[7], d[8], dp, s))
    return rc;
  VJ_HIP(hipMemcpyAsync(c, d[4], b, hipMemcpyDeviceToHost, s));
  VJ_HIP(hipMemcpyAsync(cat, d[5], b, hipMemcpyDeviceToHost, s));
  VJ_HIP(hipMemcpyAsync(sat, d[6], b, hipMemcpyDeviceToHost, s));
  VJ_HIP(hipMemcpyAsync(l1, d[7], b, hipMemcpyDeviceToHost, s));
  VJ_HIP(hipMemcpyAsync(l2, d[8], b, hipMemcpyDeviceToHost, s));
  VJ_HIP(hipMemcpyAsync(p, dp, (size_t)(nn / Lim<T>::VL) * sizeof(unsigned), hipMemcpyDeviceToHost, s));
  VJ_HIP(hipStreamSynchronize(s));
  return 0;
}

} // namespace

extern "C" {

vj_fint dbjac2_(const vj_fint *n, const double *a11, const double *a22, const double *a21, double *c, double *at, double *l1, double *l2, unsigned *p)
{
  return host_rbjac2<double>(n, a11, a22, a21, c, at, l1, l2, p);
}
vj_fint sbjac2_(const vj_fint *n, const float *a11, const float *a22, const float *a21, float *c, float *at, float *l1, float *l2, unsigned *p)
{
  return host_rbjac2<float>(n, a11, a22, a21, c, at, l1, l2, p);
}
vj_fint zbjac2_(const vj_fint *n, const double *a11, const double *a22, const double *a21r, const double *a21i, double *c, double *cat, double *sat, double *l1, double *l2, unsigned *p)
{
  return host_cbjac2<double>(n, a11, a22, a21r, a21i, c, cat, sat, l1, l2, p);
}
vj_fint cbjac2_(const vj_fint *n, const float *a11, const float *a22, const float *a21r, const float *a21i, float *c, float *cat, float *sat, float *l1, float *l2, unsigned *p)
{
  return host_cbjac2<float>(n, a11, a22, a21r, a21i, c, cat, sat, l1, l2, p);
}

vj_fint vjh_dbjac2(vj_fint n, const double *a11, const double *a22, const double *a21, double *c, double *at, double *l1, double *l2, unsigned *p, void *stream)
{
  return launch_rbjac2<double>((long long)n, a11, a22, a21, c, at, l1, l2, p, (hipStream_t)stream);
}
vj_fint vjh_sbjac2(vj_fint n, const float *a11, const float *a22, const float *a21, float *c, float *at, float *l1, float *l2, unsigned *p, void *stream)
{
  return launch_rbjac2<float>((long long)n, a11, a22, a21, c, at, l1, l2, p, (hipStream_t)stream);
}
vj_fint vjh_zbjac2(vj_fint n, const double *a11, const double *a22, const double *a21r, const double *a21i, double *c, double *cat, double *sat, double *l1, double *l2, unsigned *p, void *stream)
{
  return launch_cbjac2<double>((long long)n, a11, a22, a21r, a21i, c, cat, sat, l1, l2, p, (hipStream_t)stream);
}
vj_fint vjh_cbjac2(vj_fint n, const float *a11, const float *a22, const float *a21r, const float *a21i, float *c, float *cat, float *sat, float *l1, float *l2, unsigned *p, void *stream)
{
  return launch_cbjac2<float>((long long)n, a11, a22, a21r, a21i, c, cat, sat, l1, l2, p, (hipStream_t)stream);
}

const char *vjh_last_error(void) { return vj::last_error(); }
int vjh_abi_version(void) { return VJH_ABI_VERSION; }

} // extern "C"
