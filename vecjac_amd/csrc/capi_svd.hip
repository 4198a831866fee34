// C-ABI entry points of the plain one-sided Jacobi SVD.
//   dvjsvd_/zvjsvd_/svjsvd_/cvjsvd_: the reference's prototypes (dvjsvd.h:10,
//     zvjsvd.h:10, svjsvd.h:10, cvjsvd.h:10) on host arrays, with the
//     reference's argument checks in the reference's order;
//   vjh_?vjsvd: the same driver on device arrays and a device step table.
#include "svd_common.hpp"
#include "vecjac.h"

#include <atomic>
#include <mutex>
#include <stdint.h>
#include <stdio.h>
#include <string.h>
#include <vector>
#ifdef __x86_64__
#include <xmmintrin.h>
#endif

using namespace vj;

namespace {

inline bool fpenv_bad()
{
#ifdef __x86_64__
  return ((_mm_getcsr() & 0xFFC0u) != 0x1F80u); // vec.h:19-25
#else
  return false;
#endif
}
inline bool misaligned(const void *p) { return ((uintptr_t)p & 0x3Fu) != 0u; }

// Every row of the table must be a permutation of 0..n-1: the pairs of a
// step are then disjoint and the column updates race-free (the reference
// relies on the same invariant of its strategy tables).
bool table_ok(const unsigned *js, unsigned stp, unsigned long long n)
{
  std::vector<unsigned> seen((size_t)n, 0u);
  for (unsigned s = 0; s < stp; ++s) {
    const unsigned *r = js + (size_t)s * n;
    for (unsigned long long i = 0; i < n; ++i) {
      if (r[i] >= n || seen[r[i]] == s + 1u)
        return false;
      seen[r[i]] = s + 1u;
    }
  }
  return true;
}

__global__ void k_table_check(const unsigned *js, unsigned n, int *bad)
{
  extern __shared__ unsigned mark[];
  const unsigned *r = js + (size_t)blockIdx.x * n;
  for (unsigned i = threadIdx.x; i < (n + 31u) / 32u; i += blockDim.x)
    mark[i] = 0u;
  __syncthreads();
  for (unsigned i = threadIdx.x; i < n; i += blockDim.x) {
    const unsigned v = r[i];
    if (v >= n || (atomicOr(&mark[v >> 5], 1u << (v & 31u)) & (1u << (v & 31u))))
      atomicOr(bad, 1);
  }
}

// JTRACE-compatible trace (the reference's compile-time JTRACE output,
// dvjsvd.c:70-635, zvjsvd.c:83-692): switched on at run time with
// vjh_set_jtrace(1); the host entry points then take the file name from
// `work`, as the reference's JTRACE builds do.  The lines are the
// reference's (initial max, initial scaling, every rescaling with the step
// it precedes, one "sweep transformations" line per sweep, the final shift
// and bound).  The fused GPU step has no separate norm/dot/assembly/EVD/
// rotation phases: Tn..Te are written as 0 and Tr holds the sweeps' GPU time.
std::atomic<int> g_jtrace{0};

// The reference's compile-time USE_SECANTS form (vecdef.h:48-52,97-101,
// dvjsvd.c:411-415), switched at run time with vjh_set_secants: the pair
// decisions take secants from the 2x2 EVD and the rotations divide by them.
std::atomic<int> g_secants{0};

void write_trace(FILE *f, const PlainRun &r, int rc, bool input_error)
{
  (void)fprintf(f, "M=");
  if (input_error)
    return;
  (void)fprintf(f, "%#.17e\n", r.MG0);
  if (r.zero)
    return;
  (void)fprintf(f, "eM=%d, sR=%d, sN=%d, M=%#.17e\n", r.eM0, r.sR0, r.sN0, r.MG_scaled);
  unsigned e = 0u;
  for (unsigned sw = 0u; sw < 64u; ++sw) {
    for (; e < r.nev && e < 32u && r.ev[e].sw == sw; ++e)
      (void)fprintf(f, "sweep=%u, step=%u, eM=%d, sR=%d, sN=%d, M=%#.17e\n", r.ev[e].sw, r.ev[e].st, r.ev[e].eM,
                    r.ev[e].sR, r.ev[e].sN, r.ev[e].M);
    if (r.status && (r.err_sw == sw)) {
      (void)fprintf(f, "sweep=%u, step=%u\n", r.err_sw, r.err_st);
      return;
    }
    if (sw >= r.launched_sweeps)
      break;
    (void)fprintf(f, "%3u %llu\n", sw, r.hist[sw]);
    if (!r.hist[sw] || (sw + 1u > r.sweeps))
      break;
  }
  (void)rc;
  (void)fprintf(f, "sT=%d, M=%#.17e\n", r.sT, r.MG_final);
  (void)fprintf(f, "Tn=%15.9Lf, Tp=%15.9Lf, Ta=%15.9Lf, Te=%15.9Lf, Tr=%15.9Lf\n", 0.0L, 0.0L, 0.0L, 0.0L,
                (long double)r.steps_ms / 1000.0L);
}

template <typename T, bool CPLX>
int host_vjsvd(const vj_fnat *m, const vj_fnat *n, T *Gr, const vj_fnat *ldGr, T *Gi, const vj_fnat *ldGi,
               T *Vr, const vj_fnat *ldVr, T *Vi, const vj_fnat *ldVi, T *eS, T *fS, const unsigned *js,
               const unsigned *stp, const unsigned *swp, T *work, unsigned *iwork)
{
  constexpr unsigned VL = (unsigned)Lim<T>::VL;
  const vj_fnat n_2 = (*n >> 1u);
  if (fpenv_bad())
    return CPLX ? -18 : -14;
  if (!*n)
    return 0;
  if (*m < *n) return -1;
  if (*m & (VL - 1u)) return -1;
  if (*n & 1u) return -2;
  if (n_2 & (VL - 1u)) return -2;
  if (misaligned(Gr)) return -3;
  if (*ldGr < *m) return -4;
  if (*ldGr & (VL - 1u)) return -4;
  if (CPLX) {
    if (misaligned(Gi)) return -5;
    if (*ldGi < *m) return -6;
    if (*ldGi & (VL - 1u)) return -6;
    if (misaligned(Vr)) return -7;
    if (*ldVr < *n) return -8;
    if (*ldVr & (VL - 1u)) return -8;
    if (misaligned(Vi)) return -9;
    if (*ldVi < *n) return -10;
    if (*ldVi & (VL - 1u)) return -10;
    if (misaligned(work)) return -16;
  }
  else {
    if (misaligned(Vr)) return -5;
    if (*ldVr < *n) return -6;
    if (*ldVr & (VL - 1u)) return -6;
    if (misaligned(work)) return -12;
  }
  if (!table_ok(js, *stp, *n))
    return set_error(CPLX ? -13 : -9, "step table rows must be permutations of 0..n-1");

  const size_t M = (size_t)*m, N = (size_t)*n;
  hipStream_t s = lib_stream();
  Staging sg(s);
  PlainArgs<T> a{};
  a.m = (long long)M;
  a.n = (long long)N;
  a.ldG = (long long)M;
  a.ldV = (long long)N;
  a.stp = *stp;
  a.swp = *swp;
  a.vpreset = (*iwork != 0u);
  a.sec = g_secants.load();
  unsigned *djs = nullptr;
  if (int rc = sg.alloc(&a.Gr, M * N)) return rc;
  if (int rc = sg.alloc(&a.Vr, N * N)) return rc;
  if (CPLX) {
    if (int rc = sg.alloc(&a.Gi, M * N)) return rc;
    if (int rc = sg.alloc(&a.Vi, N * N)) return rc;
  }
  if (int rc = sg.alloc(&a.eS, N)) return rc;
  if (int rc = sg.alloc(&a.fS, N)) return rc;
  if (int rc = sg.alloc(&djs, (size_t)*stp * N)) return rc;
  a.js = djs;
  VJ_HIP(hipMemcpy2DAsync(a.Gr, M * sizeof(T), Gr, (size_t)*ldGr * sizeof(T), M * sizeof(T), N, hipMemcpyHostToDevice, s));
  if (CPLX)
    VJ_HIP(hipMemcpy2DAsync(a.Gi, M * sizeof(T), Gi, (size_t)*ldGi * sizeof(T), M * sizeof(T), N, hipMemcpyHostToDevice, s));
  if (a.vpreset) {
    VJ_HIP(hipMemcpy2DAsync(a.Vr, N * sizeof(T), Vr, (size_t)*ldVr * sizeof(T), N * sizeof(T), N, hipMemcpyHostToDevice, s));
    if (CPLX)
      VJ_HIP(hipMemcpy2DAsync(a.Vi, N * sizeof(T), Vi, (size_t)*ldVi * sizeof(T), N * sizeof(T), N, hipMemcpyHostToDevice, s));
  }
  if (*stp)
    VJ_HIP(hipMemcpyAsync(djs, js, (size_t)*stp * N * sizeof(unsigned), hipMemcpyHostToDevice, s));
  const bool trace = g_jtrace.load() != 0;
  FILE *jtr = nullptr;
  if (trace && !(jtr = fopen((const char *)work, "w")))
    return CPLX ? -85 : -73; // the reference's fopen failure (dvjsvd.c:73, zvjsvd.c:85)
  PlainRun run{};
  const int rc = plain_svd<T, CPLX>(a, trace ? &run : nullptr, s);
  const int L_MG = CPLX ? 92 : 80, L_MV = CPLX ? 104 : 92;
  // on an input error the reference returns before touching V or eS/fS
  const bool early = (rc == -L_MG) || (rc == -(L_MG + 2)) || (rc == -L_MV) || (rc == (CPLX ? -17 : -13));
  if (jtr) {
    if (rc > VJ_EHIP)
      write_trace(jtr, run, rc, early);
    (void)fclose(jtr);
  }
  if (rc <= VJ_EHIP)
    return rc;
  if (!early) {
    VJ_HIP(hipMemcpy2DAsync(Gr, (size_t)*ldGr * sizeof(T), a.Gr, M * sizeof(T), M * sizeof(T), N, hipMemcpyDeviceToHost, s));
    if (CPLX)
      VJ_HIP(hipMemcpy2DAsync(Gi, (size_t)*ldGi * sizeof(T), a.Gi, M * sizeof(T), M * sizeof(T), N, hipMemcpyDeviceToHost, s));
    VJ_HIP(hipMemcpy2DAsync(Vr, (size_t)*ldVr * sizeof(T), a.Vr, N * sizeof(T), N * sizeof(T), N, hipMemcpyDeviceToHost, s));
    if (CPLX)
      VJ_HIP(hipMemcpy2DAsync(Vi, (size_t)*ldVi * sizeof(T), a.Vi, N * sizeof(T), N * sizeof(T), N, hipMemcpyDeviceToHost, s));
    VJ_HIP(hipMemcpyAsync(eS, a.eS, N * sizeof(T), hipMemcpyDeviceToHost, s));
    VJ_HIP(hipMemcpyAsync(fS, a.fS, N * sizeof(T), hipMemcpyDeviceToHost, s));
  }
  VJ_HIP(hipStreamSynchronize(s));
  return rc;
}

template <typename T, bool CPLX>
int dev_vjsvd(vjh_nat m, vjh_nat n, T *Gr, T *Gi, vjh_nat ldG, T *Vr, T *Vi, vjh_nat ldV, T *eS, T *fS,
              const unsigned *js, unsigned stp, unsigned swp, unsigned vpreset, void *stream, unsigned *stats)
{
  if (!n)
    return 0;
  if (m < n) return -1;
  if (n & 1u) return -2;
  if (ldG < m) return -4;
  if (ldV < n) return -6;
  hipStream_t s = (hipStream_t)stream;
  if (stp) {
    int *bad = nullptr;
    VJ_HIP(hipMallocAsync((void **)&bad, sizeof(int), s));
    VJ_HIP(hipMemsetAsync(bad, 0, sizeof(int), s));
    hipLaunchKernelGGL(k_table_check, dim3(stp), dim3(256), ((n + 31) / 32) * sizeof(unsigned), s, js, (unsigned)n, bad);
    int hbad = 0;
    VJ_HIP(hipMemcpyAsync(&hbad, bad, sizeof(int), hipMemcpyDeviceToHost, s));
    VJ_HIP(hipFreeAsync(bad, s));
    VJ_HIP(hipStreamSynchronize(s));
    if (hbad)
      return set_error(CPLX ? -13 : -9, "step table rows must be permutations of 0..n-1");
  }
  PlainArgs<T> a{};
  a.m = (long long)m;
  a.n = (long long)n;
  a.Gr = Gr;
  a.Gi = Gi;
  a.ldG = (long long)ldG;
  a.Vr = Vr;
  a.Vi = Vi;
  a.ldV = (long long)ldV;
  a.eS = eS;
  a.fS = fS;
  a.js = js;
  a.stp = stp;
  a.swp = swp;
  a.vpreset = (vpreset != 0u);
  a.sec = g_secants.load();
  PlainRun run{};
  const int rc = plain_svd<T, CPLX>(a, &run, s);
  if (stats && rc > VJ_EHIP) {
    // stats[0] = sweeps, [1] = launched sweeps, [2..65] = transformations per sweep
    stats[0] = run.sweeps;
    stats[1] = run.launched_sweeps;
    for (int i = 0; i < 64; ++i)
      stats[2 + i] = (unsigned)run.hist[i];
  }
  return rc;
}

// ------------------------------------------------------ allocation-free form
// vjh_dvjsvd_noalloc keeps the scratch of one plain FP64 SVD (the control
// block, the table-check flag and the 2 * (n/2) per-pair slots) in this
// file-scope device array: one copy per GPU, held until the process ends.
// In host code `scratch` names a host-side stand-in; the device address of
// each GPU's copy comes from hipGetSymbolAddress on the first call on that
// device and is used for every kernel argument and copy.
constexpr size_t NA_STATE_OFF = 0, NA_FLAG_OFF = 8192, NA_SLOT_OFF = 8448;
constexpr size_t NA_BYTES = NA_SLOT_OFF + (size_t)VJH_NOALLOC_MAX_N * sizeof(PlainSlot);
static_assert(sizeof(PlainState<double>) <= NA_FLAG_OFF, "control block outgrew its scratch region");
static_assert(NA_SLOT_OFF % alignof(PlainSlot) == 0, "slot region misaligned");
static __device__ __align__(256) unsigned char scratch[NA_BYTES];

constexpr int NA_MAX_DEV = 64;
std::mutex g_na_mu[NA_MAX_DEV];            // one call at a time per GPU
unsigned char *g_na_ptr[NA_MAX_DEV] = {};  // device address of each GPU's scratch

int noalloc_dvjsvd(vjh_nat m, vjh_nat n, double *G, vjh_nat ldG, double *V, vjh_nat ldV, double *eS, double *fS,
                   const unsigned *js, unsigned stp, unsigned swp, unsigned vpreset, void *stream, unsigned *stats)
{
  if (!n)
    return 0;
  if (m < n) return -1;
  if (n & 1u) return -2;
  if (n > (vjh_nat)VJH_NOALLOC_MAX_N)
    return set_error(-2, "n = %llu exceeds the %d columns the allocation-free scratch holds",
                     (unsigned long long)n, VJH_NOALLOC_MAX_N);
  if (ldG < m) return -4;
  if (ldV < n) return -6;
  int dev = -1;
  VJ_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= NA_MAX_DEV)
    return set_error(VJ_EHIP, "device %d out of range", dev);
  std::lock_guard<std::mutex> lock(g_na_mu[dev]);
  if (!g_na_ptr[dev]) {
    void *p = nullptr;
    VJ_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    g_na_ptr[dev] = (unsigned char *)p;
  }
  unsigned char *const base = g_na_ptr[dev];
  hipStream_t s = (hipStream_t)stream;
  if (stp) {
    int *bad = (int *)(base + NA_FLAG_OFF);
    if (int rc = zero_words(bad, sizeof(int), s))
      return rc;
    hipLaunchKernelGGL(k_table_check, dim3(stp), dim3(256), ((n + 31) / 32) * sizeof(unsigned), s, js, (unsigned)n, bad);
    if (int rc = check_launch())
      return rc;
    int hbad = 0;
    VJ_HIP(hipMemcpyAsync(&hbad, bad, sizeof(int), hipMemcpyDeviceToHost, s));
    VJ_HIP(hipStreamSynchronize(s));
    if (hbad)
      return set_error(-9, "step table rows must be permutations of 0..n-1");
  }
  PlainArgs<double> a{};
  a.m = (long long)m;
  a.n = (long long)n;
  a.Gr = G;
  a.ldG = (long long)ldG;
  a.Vr = V;
  a.ldV = (long long)ldV;
  a.eS = eS;
  a.fS = fS;
  a.js = js;
  a.stp = stp;
  a.swp = swp;
  a.vpreset = (vpreset != 0u);
  a.sec = g_secants.load();
  PlainWork<double> w{(PlainState<double> *)(base + NA_STATE_OFF), (PlainSlot *)(base + NA_SLOT_OFF)};
  PlainRun run{};
  const int rc = plain_svd<double, false>(a, &run, s, &w);
  if (stats && rc > VJ_EHIP) {
    stats[0] = run.sweeps;
    stats[1] = run.launched_sweeps;
    for (int i = 0; i < 64; ++i)
      stats[2 + i] = (unsigned)run.hist[i];
  }
  return rc;
}

} // namespace

extern "C" {

vjh_int vjh_dvjsvd_noalloc(vjh_nat m, vjh_nat n, double *G, vjh_nat ldG, double *V, vjh_nat ldV, double *eS,
                           double *fS, const unsigned *js, unsigned stp, unsigned swp, unsigned vpreset, void *stream,
                           unsigned *stats)
{
  return noalloc_dvjsvd(m, n, G, ldG, V, ldV, eS, fS, js, stp, swp, vpreset, stream, stats);
}

void vjh_set_jtrace(int on) { g_jtrace.store(on ? 1 : 0); }

void vjh_set_secants(int on) { g_secants.store(on ? 1 : 0); }

int vjh_get_secants(void) { return g_secants.load(); }

vj_fint dvjsvd_(const vj_fnat *m, const vj_fnat *n, double *G, const vj_fnat *ldG, double *V, const vj_fnat *ldV,
                double *eS, double *fS, const unsigned *js, const unsigned *stp, const unsigned *swp, double *work,
                unsigned *iwork)
{
  return host_vjsvd<double, false>(m, n, G, ldG, nullptr, ldG, V, ldV, nullptr, ldV, eS, fS, js, stp, swp, work, iwork);
}
vj_fint svjsvd_(const vj_fnat *m, const vj_fnat *n, float *G, const vj_fnat *ldG, float *V, const vj_fnat *ldV,
                float *eS, float *fS, const unsigned *js, const unsigned *stp, const unsigned *swp, float *work,
                unsigned *iwork)
{
  return host_vjsvd<float, false>(m, n, G, ldG, nullptr, ldG, V, ldV, nullptr, ldV, eS, fS, js, stp, swp, work, iwork);
}
vj_fint zvjsvd_(const vj_fnat *m, const vj_fnat *n, double *Gr, const vj_fnat *ldGr, double *Gi, const vj_fnat *ldGi,
                double *Vr, const vj_fnat *ldVr, double *Vi, const vj_fnat *ldVi, double *eS, double *fS,
                const unsigned *js, const unsigned *stp, const unsigned *swp, double *work, unsigned *iwork)
{
  return host_vjsvd<double, true>(m, n, Gr, ldGr, Gi, ldGi, Vr, ldVr, Vi, ldVi, eS, fS, js, stp, swp, work, iwork);
}
vj_fint cvjsvd_(const vj_fnat *m, const vj_fnat *n, float *Gr, const vj_fnat *ldGr, float *Gi, const vj_fnat *ldGi,
                float *Vr, const vj_fnat *ldVr, float *Vi, const vj_fnat *ldVi, float *eS, float *fS,
                const unsigned *js, const unsigned *stp, const unsigned *swp, float *work, unsigned *iwork)
{
  return host_vjsvd<float, true>(m, n, Gr, ldGr, Gi, ldGi, Vr, ldVr, Vi, ldVi, eS, fS, js, stp, swp, work, iwork);
}

vjh_int vjh_dvjsvd(vjh_nat m, vjh_nat n, double *G, vjh_nat ldG, double *V, vjh_nat ldV, double *eS, double *fS,
                   const unsigned *js, unsigned stp, unsigned swp, unsigned vpreset, void *stream, unsigned *stats)
{
  return dev_vjsvd<double, false>(m, n, G, nullptr, ldG, V, nullptr, ldV, eS, fS, js, stp, swp, vpreset, stream, stats);
}
vjh_int vjh_svjsvd(vjh_nat m, vjh_nat n, float *G, vjh_nat ldG, float *V, vjh_nat ldV, float *eS, float *fS,
                   const unsigned *js, unsigned stp, unsigned swp, unsigned vpreset, void *stream, unsigned *stats)
{
  return dev_vjsvd<float, false>(m, n, G, nullptr, ldG, V, nullptr, ldV, eS, fS, js, stp, swp, vpreset, stream, stats);
}
vjh_int vjh_zvjsvd(vjh_nat m, vjh_nat n, double *Gr, double *Gi, vjh_nat ldG, double *Vr, double *Vi, vjh_nat ldV,
                   double *eS, double *fS, const unsigned *js, unsigned stp, unsigned swp, unsigned vpreset,
                   void *stream, unsigned *stats)
{
  return dev_vjsvd<double, true>(m, n, Gr, Gi, ldG, Vr, Vi, ldV, eS, fS, js, stp, swp, vpreset, stream, stats);
}
vjh_int vjh_cvjsvd(vjh_nat m, vjh_nat n, float *Gr, float *Gi, vjh_nat ldG, float *Vr, float *Vi, vjh_nat ldV,
                   float *eS, float *fS, const unsigned *js, unsigned stp, unsigned swp, unsigned vpreset,
                   void *stream, unsigned *stats)
{
  return dev_vjsvd<float, true>(m, n, Gr, Gi, ldG, Vr, Vi, ldV, eS, fS, js, stp, swp, vpreset, stream, stats);
}

} // extern "C"
