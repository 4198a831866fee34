// Batched 2x2 Hermitian eigendecomposition on gfx950: one lane per problem,
// FP64 (FP32) VALU, structure-of-arrays in HBM.  For finite inputs all lanes
// take one path (the special-value branches of scalef run only in waves that
// hold an infinite or NaN input).
//
// Replaces dbjac2_/zbjac2_/sbjac2_/cbjac2_ (dbjac2.c:81-142, zbjac2.c:108-171)
// and the internal dbjac2i/zbjac2i forms (dbjac2.c:178-251, zbjac2.c:222-299).
// Every input element is read once and every output written once with
// 64-lane coalesced accesses; the L1<L2 bits are gathered with one wave
// ballot and stored as the reference's packed words (VL lanes per word).
#include "vj_internal.hpp"

namespace vj {

template <typename T, int FORM>
__global__ __launch_bounds__(256) void k_rbjac2(const long long n, const T *__restrict__ a11, const T *__restrict__ a22,
                                                 const T *__restrict__ a21, T *__restrict__ c, T *__restrict__ at,
                                                 T *__restrict__ l1, T *__restrict__ l2, unsigned *__restrict__ p)
{
  const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
  bool P = false;
  if (i < n) {
    T c_, at_, l1_, l2_;
    rjac2<T, FORM>(__builtin_nontemporal_load(a11 + i), __builtin_nontemporal_load(a22 + i),
                   __builtin_nontemporal_load(a21 + i), c_, at_, l1_, l2_, P);
    __builtin_nontemporal_store(c_, c + i);
    __builtin_nontemporal_store(at_, at + i);
    __builtin_nontemporal_store(l1_, l1 + i);
    __builtin_nontemporal_store(l2_, l2 + i);
  }
  const unsigned long long mask = __ballot(P);
  constexpr int VL = Lim<T>::VL;
  const int lane = threadIdx.x & 63;
  if (((lane % VL) == 0) && (i < n))
    p[i / VL] = (unsigned)((mask >> lane) & ((1ull << VL) - 1ull));
}

template <typename T, int FORM>
__global__ __launch_bounds__(256) void k_cbjac2(const long long n, const T *__restrict__ a11, const T *__restrict__ a22,
                                                 const T *__restrict__ a21r, const T *__restrict__ a21i,
                                                 T *__restrict__ c, T *__restrict__ cat, T *__restrict__ sat,
                                                 T *__restrict__ l1, T *__restrict__ l2, unsigned *__restrict__ p)
{
  const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
  bool P = false;
  if (i < n) {
    T c_, ca_, sa_, l1_, l2_;
    cjac2<T, FORM>(__builtin_nontemporal_load(a11 + i), __builtin_nontemporal_load(a22 + i),
                   __builtin_nontemporal_load(a21r + i), __builtin_nontemporal_load(a21i + i),
                   c_, ca_, sa_, l1_, l2_, P);
    __builtin_nontemporal_store(c_, c + i);
    __builtin_nontemporal_store(ca_, cat + i);
    __builtin_nontemporal_store(sa_, sat + i);
    __builtin_nontemporal_store(l1_, l1 + i);
    __builtin_nontemporal_store(l2_, l2 + i);
  }
  const unsigned long long mask = __ballot(P);
  constexpr int VL = Lim<T>::VL;
  const int lane = threadIdx.x & 63;
  if (((lane % VL) == 0) && (i < n))
    p[i / VL] = (unsigned)((mask >> lane) & ((1ull << VL) - 1ull));
}

// Internal forms (dbjac2i/zbjac2i, dbjac2.c:178-251, zbjac2.c:222-299): a
// block of VL lanes is solved only when the low VL bits of its p word are
// not all zero, and those bits are then replaced by the block's L1 < L2
// bits; the outputs of a skipped block are left as they are.  The VL lanes
// of a word lie in one wave, so every lane reads the word before the one
// lane that rewrites it.
template <typename T, int FORM>
__global__ __launch_bounds__(256) void k_rbjac2i(const long long n, const T *__restrict__ a11, const T *__restrict__ a22,
                                                  const T *__restrict__ a21, T *__restrict__ c, T *__restrict__ at,
                                                  T *__restrict__ l1, T *__restrict__ l2, unsigned *__restrict__ p)
{
  constexpr int VL = Lim<T>::VL;
  constexpr unsigned LO = (VL == 8) ? 0xFFu : 0xFFFFu;
  const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
  const unsigned w = (i < n) ? p[i / VL] : 0u;
  const bool on = (i < n) && (w & LO);
  bool P = false;
  if (on) {
    T c_, at_, l1_, l2_;
    rjac2<T, FORM>(a11[i], a22[i], a21[i], c_, at_, l1_, l2_, P);
    c[i] = c_;
    at[i] = at_;
    l1[i] = l1_;
    l2[i] = l2_;
  }
  const unsigned long long mask = __ballot(P);
  const int lane = threadIdx.x & 63;
  if (((lane % VL) == 0) && on)
    p[i / VL] = (w & ~LO) | (unsigned)((mask >> lane) & ((1ull << VL) - 1ull));
}

template <typename T, int FORM>
__global__ __launch_bounds__(256) void k_cbjac2i(const long long n, const T *__restrict__ a11, const T *__restrict__ a22,
                                                  const T *__restrict__ a21r, const T *__restrict__ a21i,
                                                  T *__restrict__ c, T *__restrict__ cat, T *__restrict__ sat,
                                                  T *__restrict__ l1, T *__restrict__ l2, unsigned *__restrict__ p)
{
  constexpr int VL = Lim<T>::VL;
  constexpr unsigned LO = (VL == 8) ? 0xFFu : 0xFFFFu;
  const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
  const unsigned w = (i < n) ? p[i / VL] : 0u;
  const bool on = (i < n) && (w & LO);
  bool P = false;
  if (on) {
    T c_, ca_, sa_, l1_, l2_;
    cjac2<T, FORM>(a11[i], a22[i], a21r[i], a21i[i], c_, ca_, sa_, l1_, l2_, P);
    c[i] = c_;
    cat[i] = ca_;
    sat[i] = sa_;
    l1[i] = l1_;
    l2[i] = l2_;
  }
  const unsigned long long mask = __ballot(P);
  const int lane = threadIdx.x & 63;
  if (((lane % VL) == 0) && on)
    p[i / VL] = (w & ~LO) | (unsigned)((mask >> lane) & ((1ull << VL) - 1ull));
}

// n < 0: secants, n >= 0: tangents (dbjac2.c:200-250)
template <typename T>
int launch_rbjac2i(long long n, const T *a11, const T *a22, const T *a21, T *c, T *at, T *l1, T *l2, unsigned *p,
                   hipStream_t s)
{
  const long long nn = (n < 0) ? -n : n;
  if (nn % Lim<T>::VL)
    return -1;
  if (!nn)
    return 0;
  const long long blocks = (nn + 255) / 256;
  if (n < 0)
    hipLaunchKernelGGL((k_rbjac2i<T, JAC2_SEC>), dim3((unsigned)blocks), dim3(256), 0, s, nn, a11, a22, a21, c, at, l1, l2, p);
  else
    hipLaunchKernelGGL((k_rbjac2i<T, JAC2_TAN>), dim3((unsigned)blocks), dim3(256), 0, s, nn, a11, a22, a21, c, at, l1, l2, p);
  return check_launch();
}

template <typename T>
int launch_cbjac2i(long long n, const T *a11, const T *a22, const T *a21r, const T *a21i, T *c, T *cat, T *sat, T *l1,
                   T *l2, unsigned *p, hipStream_t s)
{
  const long long nn = (n < 0) ? -n : n;
  if (nn % Lim<T>::VL)
    return -1;
  if (!nn)
    return 0;
  const long long blocks = (nn + 255) / 256;
  if (n < 0)
    hipLaunchKernelGGL((k_cbjac2i<T, JAC2_SEC>), dim3((unsigned)blocks), dim3(256), 0, s, nn, a11, a22, a21r, a21i, c, cat, sat, l1, l2, p);
  else
    hipLaunchKernelGGL((k_cbjac2i<T, JAC2_TAN>), dim3((unsigned)blocks), dim3(256), 0, s, nn, a11, a22, a21r, a21i, c, cat, sat, l1, l2, p);
  return check_launch();
}

template <typename T>
int launch_rbjac2(long long n, const T *a11, const T *a22, const T *a21, T *c, T *at, T *l1, T *l2, unsigned *p, hipStream_t s)
{
  const int form = (n < 0) ? JAC2_SIN : JAC2_TAN;
  const long long nn = (n < 0) ? -n : n;
  if (nn % Lim<T>::VL)
    return -1;
  if (!nn)
    return 0;
  const long long blocks = (nn + 255) / 256;
  if (form == JAC2_SIN)
    hipLaunchKernelGGL((k_rbjac2<T, JAC2_SIN>), dim3((unsigned)blocks), dim3(256), 0, s, nn, a11, a22, a21, c, at, l1, l2, p);
  else
    hipLaunchKernelGGL((k_rbjac2<T, JAC2_TAN>), dim3((unsigned)blocks), dim3(256), 0, s, nn, a11, a22, a21, c, at, l1, l2, p);
  return check_launch();
}

template <typename T>
int launch_cbjac2(long long n, const T *a11, const T *a22, const T *a21r, const T *a21i, T *c, T *cat, T *sat, T *l1, T *l2, unsigned *p, hipStream_t s)
{
  const int form = (n < 0) ? JAC2_SIN : JAC2_TAN;
  const long long nn = (n < 0) ? -n : n;
  if (nn % Lim<T>::VL)
    return -1;
  if (!nn)
    return 0;
  const long long blocks = (nn + 255) / 256;
  if (form == JAC2_SIN)
    hipLaunchKernelGGL((k_cbjac2<T, JAC2_SIN>), dim3((unsigned)blocks), dim3(256), 0, s, nn, a11, a22, a21r, a21i, c, cat, sat, l1, l2, p);
  else
    hipLaunchKernelGGL((k_cbjac2<T, JAC2_TAN>), dim3((unsigned)blocks), dim3(256), 0, s, nn, a11, a22, a21r, a21i, c, cat, sat, l1, l2, p);
  return check_launch();
}

template int launch_rbjac2<double>(long long, const double *, const double *, const double *, double *, double *, double *, double *, unsigned *, hipStream_t);
template int launch_rbjac2<float>(long long, const float *, const float *, const float *, float *, float *, float *, float *, unsigned *, hipStream_t);
template int launch_cbjac2<double>(long long, const double *, const double *, const double *, const double *, double *, double *, double *, double *, double *, unsigned *, hipStream_t);
template int launch_cbjac2<float>(long long, const float *, const float *, const float *, const float *, float *, float *, float *, float *, float *, unsigned *, hipStream_t);
template int launch_rbjac2i<double>(long long, const double *, const double *, const double *, double *, double *, double *, double *, unsigned *, hipStream_t);
template int launch_rbjac2i<float>(long long, const float *, const float *, const float *, float *, float *, float *, float *, unsigned *, hipStream_t);
template int launch_cbjac2i<double>(long long, const double *, const double *, const double *, const double *, double *, double *, double *, double *, double *, unsigned *, hipStream_t);
template int launch_cbjac2i<float>(long long, const float *, const float *, const float *, const float *, float *, float *, float *, float *, float *, unsigned *, hipStream_t);

} // namespace vj
