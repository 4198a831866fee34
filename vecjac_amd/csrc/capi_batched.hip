// C-ABI of the batched SVD (no reference counterpart): a plan holds the
// shape-dependent metadata on the device (descriptors and the cost order),
// so that a run is a single kernel launch on resident data and can be
// captured in a HIP graph.  The kernel computes the pairs of each step
// itself: the modulus ordering, the table pjs(4, n).
#include "svd_batched.hpp"
#include "svd_common.hpp"
#include "vecjac.h"

#include <algorithm>
#include <map>
#include <numeric>
#include <stdlib.h>
#include <vector>

using namespace vj;

namespace {

struct BPlan {
  int kind = 0;            // 0 d, 1 z, 2 s, 3 c
  int nb = 0;
  int max_m = 0;
  int nlong = 0;           // matrices with m > short_rows: first in the order, one launch
  int max_m_long = 0, max_m_short = 0;
  int waves = 4;           // waves per workgroup of the long-column launch
  int waves_short = 4;     // and of the short-column launch
  long long totG = 0;      // elements of the packed G (max over b of offG + m n)
  unsigned swp = 999u;
  int dev = 0;
  long long *d_meta = nullptr; // m, n, offG, offV, offS (nb each)
  int *d_order = nullptr;
  unsigned *d_queue = nullptr; // work counters of the two launches
  unsigned long long *d_evals = nullptr; // pair evaluations per matrix (last run)
};

__global__ void k_zero_u32(unsigned *p)
{
  p[0] = 0u;
  p[1] = 0u;
}

int pick_waves(const std::vector<long long> &n)
{
  std::vector<long long> s(n);
  std::nth_element(s.begin(), s.begin() + s.size() / 2, s.end());
  const long long med = s[s.size() / 2];
  return (med <= 24) ? 1 : ((med <= 48) ? 2 : 4);
}

// Eight waves per workgroup leave 256 registers per lane; plans whose
// (double-buffered) column pairs would spill at that size run four waves.
int cap_waves(int kind, int max_m, int waves)
{
  const int e = (max_m <= 64) ? 1 : (max_m <= 128) ? 2 : (max_m <= 256) ? 4 : (max_m <= 512) ? 8 : 16;
  const int bytes = e * ((kind == 0 || kind == 1) ? 8 : 4) * ((kind & 1) ? 2 : 1);
  const int limit = (kind == 1) ? 16 : ((kind == 3) ? 32 : 64);
  return (waves > 4 && bytes > limit) ? 4 : waves;
}

// mid, if not null, is recorded between the two launches (autotuner timing)
template <typename T, bool CPLX>
int run_plan(const BPlan *P, void *Gr, void *Gi, void *Vr, void *Vi, void *eS, void *fS, vj_fint *info,
             uint64_t *transforms, hipStream_t s, hipEvent_t mid = nullptr)
{
  BatchArgs<T> a{};
  const long long nb = P->nb;
  a.nb = P->nb;
  a.m = P->d_meta;
  a.n = P->d_meta + nb;
  a.offG = P->d_meta + 2 * nb;
  a.offV = P->d_meta + 3 * nb;
  a.offS = P->d_meta + 4 * nb;
  a.order = P->d_order;
  a.Gr = (T *)Gr;
  a.Gi = (T *)Gi;
  a.Vr = (T *)Vr;
  a.Vi = (T *)Vi;
  a.eS = (T *)eS;
  a.fS = (T *)fS;
  a.info = (long long *)info;
  a.transforms = (unsigned long long *)transforms;
  a.evals = P->d_evals;
  a.swp = P->swp;
  // a fill kernel, not hipMemsetAsync, so that the run can be graph-captured
  hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(1), 0, s, P->d_queue);
  // the long columns first (the costly matrices start first), then the
  // short ones in the lane-group form, each launch with its own queue
  a.nb = P->nlong;
  a.queue = P->d_queue;
  if (int rc = batched_svd<T, CPLX>(a, P->waves, P->max_m_long, 0, false, s))
    return rc;
  if (mid && hipEventRecord(mid, s) != hipSuccess)
    return set_error(VJ_EHIP, "event record failed");
  a.nb = P->nb - P->nlong;
  a.order = P->d_order + P->nlong;
  a.queue = P->d_queue + 1;
  return batched_svd<T, CPLX>(a, P->waves_short, P->max_m_short, 0, true, s);
}

int run_any(const BPlan *P, void *Gr, void *Gi, void *Vr, void *Vi, void *eS, void *fS, vj_fint *info,
            uint64_t *transforms, hipStream_t s, hipEvent_t mid)
{
  switch (P->kind) {
  case 0: return run_plan<double, false>(P, Gr, nullptr, Vr, nullptr, eS, fS, info, transforms, s, mid);
  case 1: return run_plan<double, true>(P, Gr, Gi, Vr, Vi, eS, fS, info, transforms, s, mid);
  case 2: return run_plan<float, false>(P, Gr, nullptr, Vr, nullptr, eS, fS, info, transforms, s, mid);
  default: return run_plan<float, true>(P, Gr, Gi, Vr, Vi, eS, fS, info, transforms, s, mid);
  }
}

} // namespace

extern "C" {

void *vjh_bplan_create(int kind, vj_fint nb, const vj_fint *m, const vj_fint *n, const vj_fint *offG,
                       const vj_fint *offV, const vj_fint *offS, unsigned swp, int waves, vj_fint *status)
{
  vj_fint dummy = 0;
  vj_fint *st = status ? status : &dummy;
  *st = 0;
  if (kind < 0 || kind > 3 || nb < 0) {
    *st = set_error(-1, "bad kind or batch size");
    return nullptr;
  }
  std::vector<long long> hm(m, m + nb), hn(n, n + nb);
  int max_m = 0;
  for (vj_fint b = 0; b < nb; ++b) {
    if (hn[b] < 2 || (hn[b] & 1) || hn[b] > 128 || hm[b] < hn[b] || hm[b] > 1024 || offG[b] < 0 || offV[b] < 0 ||
        offS[b] < 0) {
      *st = set_error(-2, "matrix %lld: need even 2 <= n <= 128, n <= m <= 1024 and non-negative offsets (m=%lld n=%lld)",
                      (long long)b, hm[b], hn[b]);
      return nullptr;
    }
    max_m = std::max(max_m, (int)hm[b]);
  }
  long long totG = 0;
  for (vj_fint b = 0; b < nb; ++b)
    totG = std::max(totG, (long long)offG[b] + hm[b] * hn[b]);
  BPlan *P = new BPlan();
  P->kind = kind;
  P->nb = (int)nb;
  P->max_m = max_m;
  P->totG = totG;
  P->swp = swp;
  const int w0 = (waves > 0) ? waves : (nb ? pick_waves(hn) : 4);
  P->waves = cap_waves(kind, max_m, w0);
  (void)hipGetDevice(&P->dev);
  // processing order: the long-column class, then the short one, each by
  // decreasing m n^2 (longest first)
  const long long sr = short_rows((kind & 1) != 0);
  std::vector<int> order((size_t)nb);
  std::iota(order.begin(), order.end(), 0);
  std::stable_sort(order.begin(), order.end(), [&](int x, int y) {
    const bool lx = hm[x] > sr, ly = hm[y] > sr;
    if (lx != ly)
      return lx;
    return hm[x] * hn[x] * hn[x] > hm[y] * hn[y] * hn[y];
  });
  for (vj_fint b = 0; b < nb; ++b) {
    if (hm[b] > sr) {
      ++P->nlong;
      P->max_m_long = std::max(P->max_m_long, (int)hm[b]);
    }
    else
      P->max_m_short = std::max(P->max_m_short, (int)hm[b]);
  }
  P->waves_short = cap_waves(kind, P->max_m_short, w0);
  std::vector<long long> meta;
  meta.reserve((size_t)(5 * nb));
  meta.insert(meta.end(), hm.begin(), hm.end());
  meta.insert(meta.end(), hn.begin(), hn.end());
  meta.insert(meta.end(), offG, offG + nb);
  meta.insert(meta.end(), offV, offV + nb);
  meta.insert(meta.end(), offS, offS + nb);
  hipError_t e = hipSuccess;
  if (e == hipSuccess) e = hipMalloc((void **)&P->d_meta, std::max<size_t>(1, meta.size()) * sizeof(long long));
  if (e == hipSuccess) e = hipMalloc((void **)&P->d_order, std::max<size_t>(1, order.size()) * sizeof(int));
  if (e == hipSuccess) e = hipMalloc((void **)&P->d_queue, 2 * sizeof(unsigned));
  if (e == hipSuccess) e = hipMalloc((void **)&P->d_evals, std::max<size_t>(1, (size_t)nb) * sizeof(unsigned long long));
  if (e == hipSuccess && !meta.empty()) e = hipMemcpy(P->d_meta, meta.data(), meta.size() * sizeof(long long), hipMemcpyHostToDevice);
  if (e == hipSuccess && !order.empty()) e = hipMemcpy(P->d_order, order.data(), order.size() * sizeof(int), hipMemcpyHostToDevice);
  if (e != hipSuccess) {
    *st = set_error(VJ_EHIP - (int)e, "plan allocation failed: %s", hipGetErrorString(e));
    (void)hipFree(P->d_meta);
    (void)hipFree(P->d_order);
    (void)hipFree(P->d_queue);
    (void)hipFree(P->d_evals);
    delete P;
    return nullptr;
  }
  return P;
}

vj_fint vjh_bplan_run(void *plan, void *Gr, void *Gi, void *Vr, void *Vi, void *eS, void *fS, vj_fint *info,
                      uint64_t *transforms, void *stream)
{
  const BPlan *P = (const BPlan *)plan;
  if (!P)
    return set_error(-1, "null plan");
  return run_any(P, Gr, Gi, Vr, Vi, eS, fS, info, transforms, (hipStream_t)stream, nullptr);
}

int vjh_bplan_waves(const void *plan) { return plan ? ((const BPlan *)plan)->waves : 0; }

vj_fint vjh_bplan_evals(const void *plan, uint64_t *evals, void *stream)
{
  const BPlan *P = (const BPlan *)plan;
  if (!P || !evals)
    return set_error(-1, "null plan or output");
  hipStream_t s = (hipStream_t)stream;
  if (P->nb)
    VJ_HIP(hipMemcpyAsync(evals, P->d_evals, (size_t)P->nb * sizeof(uint64_t), hipMemcpyDeviceToHost, s));
  VJ_HIP(hipStreamSynchronize(s));
  return 0;
}

vj_fint vjh_bplan_autotune(void *plan, const void *G, const void *Gi, void *Gw, void *Gwi, void *V, void *Vi, void *eS,
                           void *fS, vj_fint *info, float *ms, void *stream)
{
  BPlan *P = (BPlan *)plan;
  if (!P)
    return set_error(-1, "null plan");
  hipStream_t s = (hipStream_t)stream;
  const bool cplx = (P->kind & 1) != 0;
  const size_t bytes = (size_t)P->totG * ((P->kind <= 1) ? sizeof(double) : sizeof(float));
  hipEvent_t e0 = nullptr, e1 = nullptr;
  VJ_HIP(hipEventCreate(&e0));
  VJ_HIP(hipEventCreate(&e1));
  static const int cand[4] = {1, 2, 4, 8};
  // each launch is timed on its own (events before, between and after the
  // two launches) and keeps its own fastest count
  hipEvent_t em = nullptr;
  VJ_HIP(hipEventCreate(&em));
  int best = P->waves, best_s = P->waves_short;
  float best_ms = -1.0f, best_ms_s = -1.0f;
  vj_fint rc = 0;
  for (int k = 0; k < 4 && !rc; ++k) {
    const int w = cap_waves(P->kind, P->max_m, cand[k]);
    const int ws = cap_waves(P->kind, P->max_m_short, cand[k]);
    float t = -1.0f, tl = -1.0f, ts = -1.0f;
    if ((w != cand[k]) && (ws != cand[k])) { // both capped: the same launches as a smaller count
      if (ms)
        ms[k] = t;
      continue;
    }
    P->waves = w;
    P->waves_short = ws;
    if (bytes) {
      if ((rc = (vj_fint)hipMemcpyAsync(Gw, G, bytes, hipMemcpyDeviceToDevice, s)))
        rc = set_error(VJ_EHIP - (int)rc, "autotune: copy failed");
      if (!rc && cplx && (rc = (vj_fint)hipMemcpyAsync(Gwi, Gi, bytes, hipMemcpyDeviceToDevice, s)))
        rc = set_error(VJ_EHIP - (int)rc, "autotune: copy failed");
    }
    if (!rc && (rc = (vj_fint)hipEventRecord(e0, s)))
      rc = set_error(VJ_EHIP - (int)rc, "autotune: event failed");
    if (!rc)
      rc = run_any(P, Gw, Gwi, V, Vi, eS, fS, info, nullptr, s, em);
    if (!rc && (rc = (vj_fint)hipEventRecord(e1, s)))
      rc = set_error(VJ_EHIP - (int)rc, "autotune: event failed");
    if (!rc && (rc = (vj_fint)hipEventSynchronize(e1)))
      rc = set_error(VJ_EHIP - (int)rc, "autotune: synchronisation failed");
    if (!rc && ((rc = (vj_fint)hipEventElapsedTime(&tl, e0, em)) || (rc = (vj_fint)hipEventElapsedTime(&ts, em, e1))))
      rc = set_error(VJ_EHIP - (int)rc, "autotune: timing failed");
    if (!rc) {
      t = tl + ts;
      // a capped count is the same launch as a smaller one: keep the first
      if ((w == cand[k]) && (best_ms < 0.0f || tl < best_ms)) {
        best_ms = tl;
        best = w;
      }
      if ((ws == cand[k]) && (best_ms_s < 0.0f || ts < best_ms_s)) {
        best_ms_s = ts;
        best_s = ws;
      }
    }
    if (ms)
      ms[k] = t;
  }
  (void)hipEventDestroy(em);
  P->waves_short = best_s;
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  P->waves = best;
  return rc ? rc : (vj_fint)best;
}

void vjh_bplan_destroy(void *plan)
{
  BPlan *P = (BPlan *)plan;
  if (!P)
    return;
  (void)hipFree(P->d_meta);
  (void)hipFree(P->d_order);
  (void)hipFree(P->d_queue);
  (void)hipFree(P->d_evals);
  delete P;
}

vj_fint vjh_dbvjsvd(vj_fint nb, const vj_fint *m, const vj_fint *n, const vj_fint *offG, const vj_fint *offV,
                    const vj_fint *offS, double *G, double *V, double *eS, double *fS, vj_fint *info,
                    uint64_t *transforms, unsigned swp, int waves, void *stream)
{
  vj_fint st = 0;
  void *P = vjh_bplan_create(0, nb, m, n, offG, offV, offS, swp, waves, &st);
  if (!P)
    return st;
  const vj_fint rc = vjh_bplan_run(P, G, nullptr, V, nullptr, eS, fS, info, transforms, stream);
  (void)hipStreamSynchronize((hipStream_t)stream);
  vjh_bplan_destroy(P);
  return rc;
}

} // extern "C"
