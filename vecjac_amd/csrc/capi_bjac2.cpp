// C-ABI entry points of the batched 2x2 Hermitian EVD.
//   dbjac2_/zbjac2_/sbjac2_/cbjac2_: the reference's prototypes (dbjac2.h:6,
//     zbjac2.h:6, sbjac2.h:6, cbjac2.h:6), host arrays in and out;
//   vjh_?bjac2: the same computation on device arrays, enqueued on a stream.
//   dbjac2i/zbjac2i/sbjac2i/cbjac2i: the internal forms (dbjac2.h:9,
//     dbjac2.c:178-251): n<0 secants, n>=0 tangents, a block of VL lanes
//     solved only when the low VL bits of its p word are set, those bits
//     replaced by the block's L1<L2 bits (p is read and written).
// The sign of n selects sines (n<0) or tangents (n>0), as in the reference.
// |n| must be a multiple of the reference's lane count (8 for FP64, 16 for
// FP32) because p[] packs that many L1<L2 bits per word; otherwise -1 (the
// reference's own argument check, dbjac2.c:86-87).  The host entry points
// return 1 on success, as the reference's release build (OpenMP) does
// (dbjac2.c:113, 132); the device entry points return 0.
#include "capi_common.hpp"
#include "vecjac.h"

using namespace vj;

namespace {

constexpr int HOST_OK = 1; // the reference's release (OpenMP) build returns 1

template <typename T, bool INTERNAL>
int host_rbjac2(const vj_fint *n, const T *a11, const T *a22, const T *a21, T *c, T *at, T *l1, T *l2, unsigned *p)
{
  const long long nn = (*n < 0) ? -(long long)*n : (long long)*n;
  if (nn % Lim<T>::VL)
    return -1;
  if (!nn)
    return HOST_OK;
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
  if (INTERNAL) {
    // a skipped block keeps its outputs and p word: stage them in as well
    VJ_HIP(hipMemcpyAsync(dp, p, (size_t)(nn / Lim<T>::VL) * sizeof(unsigned), hipMemcpyHostToDevice, s));
    VJ_HIP(hipMemcpyAsync(d[3], c, b, hipMemcpyHostToDevice, s));
    VJ_HIP(hipMemcpyAsync(d[4], at, b, hipMemcpyHostToDevice, s));
    VJ_HIP(hipMemcpyAsync(d[5], l1, b, hipMemcpyHostToDevice, s));
    VJ_HIP(hipMemcpyAsync(d[6], l2, b, hipMemcpyHostToDevice, s));
    if (int rc = launch_rbjac2i<T>((long long)*n, d[0], d[1], d[2], d[3], d[4], d[5], d[6], dp, s))
      return rc;
  }
  else if (int rc = launch_rbjac2<T>((long long)*n, d[0], d[1], d[2], d[3], d[4], d[5], d[6], dp, s))
    return rc;
  VJ_HIP(hipMemcpyAsync(c, d[3], b, hipMemcpyDeviceToHost, s));
  VJ_HIP(hipMemcpyAsync(at, d[4], b, hipMemcpyDeviceToHost, s));
  VJ_HIP(hipMemcpyAsync(l1, d[5], b, hipMemcpyDeviceToHost, s));
  VJ_HIP(hipMemcpyAsync(l2, d[6], b, hipMemcpyDeviceToHost, s));
  VJ_HIP(hipMemcpyAsync(p, dp, (size_t)(nn / Lim<T>::VL) * sizeof(unsigned), hipMemcpyDeviceToHost, s));
  VJ_HIP(hipStreamSynchronize(s));
  return HOST_OK;
}

template <typename T, bool INTERNAL>
int host_cbjac2(const vj_fint *n, const T *a11, const T *a22, const T *a21r, const T *a21i, T *c, T *cat, T *sat, T *l1, T *l2, unsigned *p)
{
  const long long nn = (*n < 0) ? -(long long)*n : (long long)*n;
  if (nn % Lim<T>::VL)
    return -1;
  if (!nn)
    return HOST_OK;
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
  if (INTERNAL) {
    VJ_HIP(hipMemcpyAsync(dp, p, (size_t)(nn / Lim<T>::VL) * sizeof(unsigned), hipMemcpyHostToDevice, s));
    VJ_HIP(hipMemcpyAsync(d[4], c, b, hipMemcpyHostToDevice, s));
    VJ_HIP(hipMemcpyAsync(d[5], cat, b, hipMemcpyHostToDevice, s));
    VJ_HIP(hipMemcpyAsync(d[6], sat, b, hipMemcpyHostToDevice, s));
    VJ_HIP(hipMemcpyAsync(d[7], l1, b, hipMemcpyHostToDevice, s));
    VJ_HIP(hipMemcpyAsync(d[8], l2, b, hipMemcpyHostToDevice, s));
    if (int rc = launch_cbjac2i<T>((long long)*n, d[0], d[1], d[2], d[3], d[4], d[5], d[6], d[7], d[8], dp, s))
      return rc;
  }
  else if (int rc = launch_cbjac2<T>((long long)*n, d[0], d[1], d[2], d[3], d[4], d[5], d[6], d[7], d[8], dp, s))
    return rc;
  VJ_HIP(hipMemcpyAsync(c, d[4], b, hipMemcpyDeviceToHost, s));
  VJ_HIP(hipMemcpyAsync(cat, d[5], b, hipMemcpyDeviceToHost, s));
  VJ_HIP(hipMemcpyAsync(sat, d[6], b, hipMemcpyDeviceToHost, s));
  VJ_HIP(hipMemcpyAsync(l1, d[7], b, hipMemcpyDeviceToHost, s));
  VJ_HIP(hipMemcpyAsync(l2, d[8], b, hipMemcpyDeviceToHost, s));
  VJ_HIP(hipMemcpyAsync(p, dp, (size_t)(nn / Lim<T>::VL) * sizeof(unsigned), hipMemcpyDeviceToHost, s));
  VJ_HIP(hipStreamSynchronize(s));
  return HOST_OK;
}

} // namespace

extern "C" {

vj_fint dbjac2_(const vj_fint *n, const double *a11, const double *a22, const double *a21, double *c, double *at, double *l1, double *l2, unsigned *p)
{
  return host_rbjac2<double, false>(n, a11, a22, a21, c, at, l1, l2, p);
}
vj_fint sbjac2_(const vj_fint *n, const float *a11, const float *a22, const float *a21, float *c, float *at, float *l1, float *l2, unsigned *p)
{
  return host_rbjac2<float, false>(n, a11, a22, a21, c, at, l1, l2, p);
}
vj_fint zbjac2_(const vj_fint *n, const double *a11, const double *a22, const double *a21r, const double *a21i, double *c, double *cat, double *sat, double *l1, double *l2, unsigned *p)
{
  return host_cbjac2<double, false>(n, a11, a22, a21r, a21i, c, cat, sat, l1, l2, p);
}
vj_fint cbjac2_(const vj_fint *n, const float *a11, const float *a22, const float *a21r, const float *a21i, float *c, float *cat, float *sat, float *l1, float *l2, unsigned *p)
{
  return host_cbjac2<float, false>(n, a11, a22, a21r, a21i, c, cat, sat, l1, l2, p);
}
vj_fint dbjac2i(const vj_fint *n, const double *a11, const double *a22, const double *a21, double *c, double *at, double *l1, double *l2, unsigned *p)
{
  return host_rbjac2<double, true>(n, a11, a22, a21, c, at, l1, l2, p);
}
vj_fint sbjac2i(const vj_fint *n, const float *a11, const float *a22, const float *a21, float *c, float *at, float *l1, float *l2, unsigned *p)
{
  return host_rbjac2<float, true>(n, a11, a22, a21, c, at, l1, l2, p);
}
vj_fint zbjac2i(const vj_fint *n, const double *a11, const double *a22, const double *a21r, const double *a21i, double *c, double *cat, double *sat, double *l1, double *l2, unsigned *p)
{
  return host_cbjac2<double, true>(n, a11, a22, a21r, a21i, c, cat, sat, l1, l2, p);
}
vj_fint cbjac2i(const vj_fint *n, const float *a11, const float *a22, const float *a21r, const float *a21i, float *c, float *cat, float *sat, float *l1, float *l2, unsigned *p)
{
  return host_cbjac2<float, true>(n, a11, a22, a21r, a21i, c, cat, sat, l1, l2, p);
}

vjh_int vjh_dbjac2(vjh_int n, const double *a11, const double *a22, const double *a21, double *c, double *at, double *l1, double *l2, unsigned *p, void *stream)
{
  return launch_rbjac2<double>((long long)n, a11, a22, a21, c, at, l1, l2, p, (hipStream_t)stream);
}
vjh_int vjh_sbjac2(vjh_int n, const float *a11, const float *a22, const float *a21, float *c, float *at, float *l1, float *l2, unsigned *p, void *stream)
{
  return launch_rbjac2<float>((long long)n, a11, a22, a21, c, at, l1, l2, p, (hipStream_t)stream);
}
vjh_int vjh_zbjac2(vjh_int n, const double *a11, const double *a22, const double *a21r, const double *a21i, double *c, double *cat, double *sat, double *l1, double *l2, unsigned *p, void *stream)
{
  return launch_cbjac2<double>((long long)n, a11, a22, a21r, a21i, c, cat, sat, l1, l2, p, (hipStream_t)stream);
}
vjh_int vjh_cbjac2(vjh_int n, const float *a11, const float *a22, const float *a21r, const float *a21i, float *c, float *cat, float *sat, float *l1, float *l2, unsigned *p, void *stream)
{
  return launch_cbjac2<float>((long long)n, a11, a22, a21r, a21i, c, cat, sat, l1, l2, p, (hipStream_t)stream);
}

vjh_int vjh_dbjac2i(vjh_int n, const double *a11, const double *a22, const double *a21, double *c, double *at, double *l1, double *l2, unsigned *p, void *stream)
{
  return launch_rbjac2i<double>((long long)n, a11, a22, a21, c, at, l1, l2, p, (hipStream_t)stream);
}
vjh_int vjh_sbjac2i(vjh_int n, const float *a11, const float *a22, const float *a21, float *c, float *at, float *l1, float *l2, unsigned *p, void *stream)
{
  return launch_rbjac2i<float>((long long)n, a11, a22, a21, c, at, l1, l2, p, (hipStream_t)stream);
}
vjh_int vjh_zbjac2i(vjh_int n, const double *a11, const double *a22, const double *a21r, const double *a21i, double *c, double *cat, double *sat, double *l1, double *l2, unsigned *p, void *stream)
{
  return launch_cbjac2i<double>((long long)n, a11, a22, a21r, a21i, c, cat, sat, l1, l2, p, (hipStream_t)stream);
}
vjh_int vjh_cbjac2i(vjh_int n, const float *a11, const float *a22, const float *a21r, const float *a21i, float *c, float *cat, float *sat, float *l1, float *l2, unsigned *p, void *stream)
{
  return launch_cbjac2i<float>((long long)n, a11, a22, a21r, a21i, c, cat, sat, l1, l2, p, (hipStream_t)stream);
}

const char *vjh_last_error(void) { return vj::last_error(); }
int vjh_abi_version(void) { return VJH_ABI_VERSION; }
int vjh_fint_bits(void) { return (int)(8 * sizeof(vj_fint)); }

} // extern "C"
