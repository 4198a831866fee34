// C-ABI of the batched SVD (no reference counterpart): a plan holds the
// shape-dependent metadata on the device (descriptors and the cost order),
// so that a run is a single kernel launch on resident data and can be
// captured in a HIP graph.  The kernel computes the pairs of each step
// itself: the block modulus ordering, the table pjs(6, n), by default, or
// the modulus ordering, pjs(4, n).
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

// A run is three launches, one per class of matrices, each class in its own
// stretch of the processing order with its own work counter:
//   0: long columns (m > short_rows), the wave form;
//   1: short columns whose G exceeds the LDS budget, the lane-group form on
//      global memory;
//   2: short columns whose G fits the budget, the lane-group form with G
//      resident in LDS (V in device memory).
constexpr int NCLASS = 3;

struct BPlan {
  int kind = 0;            // 0 d, 1 z, 2 s, 3 c
  int nb = 0;
  int max_m = 0;
  int count[NCLASS] = {};  // matrices of each class
  int max_m_c[NCLASS] = {};
  int waves_c[NCLASS] = {4, 4, 4}; // waves per workgroup of each launch
  int lg_c[NCLASS] = {4, 4, 4};    // lane-group launches: 2^lg lanes per pair
  int split_rows = 128;     // matrices with more rows take the wave form
  long long lds_budget = 0; // bytes of G up to which a short matrix is LDS-resident
  int lds_bytes = 0;       // dynamic LDS per workgroup of the LDS launch (max over its class)
  long long totG = 0;      // elements of the packed G (max over b of offG + m n)
  unsigned swp = 999u;
  int ord = 6;             // ordering of the pairs (pjs id)
  int queue = 0;           // cost model of the processing order (queue_cost)
  int vpreset = 0;         // V holds the starting matrices on entry
  int dev = 0;
  std::vector<long long> hm, hn; // the shapes, for re-partitioning
  long long *d_meta = nullptr; // m, n, offG, offV, offS (nb each)
  int *d_order = nullptr;
  unsigned *d_queue = nullptr; // work counters of the launches
  unsigned long long *d_evals = nullptr; // pair evaluations per matrix (last run)
  unsigned long long *d_xfer = nullptr;  // elements moved per matrix (last run)
};

__global__ void k_zero_u32(unsigned *p)
{
  for (int i = 0; i < NCLASS; ++i)
    p[i] = 0u;
}

// Default LDS budget: G of up to 48 KiB, so that three four-wave workgroups
// (the register limit of the lane-group form) fit a CU.
constexpr long long DEFAULT_LDS_BUDGET = 48 * 1024;

int tsize(int kind) { return (kind <= 1) ? 8 : 4; }

int class_of(const BPlan *P, long long m, long long n)
{
  const bool cplx = (P->kind & 1) != 0;
  if (m > P->split_rows)
    return 0;
  return (lds_matrix_bytes(m, n, cplx, tsize(P->kind)) <= P->lds_budget) ? 2 : 1;
}

// Estimated cost of a matrix under the plan's queue model: 0, m n^2 (the
// column traffic of a sweep); 1 and 2, n^2 (c + m / 64) with c = 8 or 16 (a
// pair evaluation costs a fixed decision stream plus a share per 64-row
// chunk, the fixed part measured at about 15 chunks' worth).
double queue_cost(const BPlan *P, long long m, long long n)
{
  const double md = (double)m, nd = (double)n;
  switch (P->queue) {
  case 1: return nd * nd * (8.0 + md / 64.0);
  case 2: return nd * nd * (16.0 + md / 64.0);
  default: return md * nd * nd;
  }
}

// Processing order (the classes in turn, each by decreasing estimated cost,
// the costliest first), class counts and row maxima; uploads the order.
int partition(BPlan *P)
{
  const long long nb = P->nb;
  std::vector<int> order((size_t)nb), cls((size_t)nb);
  for (long long b = 0; b < nb; ++b)
    cls[(size_t)b] = class_of(P, P->hm[(size_t)b], P->hn[(size_t)b]);
  std::iota(order.begin(), order.end(), 0);
  std::stable_sort(order.begin(), order.end(), [&](int x, int y) {
    if (cls[(size_t)x] != cls[(size_t)y])
      return cls[(size_t)x] < cls[(size_t)y];
    return queue_cost(P, P->hm[(size_t)x], P->hn[(size_t)x]) > queue_cost(P, P->hm[(size_t)y], P->hn[(size_t)y]);
  });
  for (int c = 0; c < NCLASS; ++c)
    P->count[c] = P->max_m_c[c] = 0;
  long long lds = 0;
  const bool cplx = (P->kind & 1) != 0;
  for (long long b = 0; b < nb; ++b) {
    const int c = cls[(size_t)b];
    ++P->count[c];
    P->max_m_c[c] = std::max(P->max_m_c[c], (int)P->hm[(size_t)b]);
    if (c == 2)
      lds = std::max(lds, lds_matrix_bytes(P->hm[(size_t)b], P->hn[(size_t)b], cplx, tsize(P->kind)));
  }
  P->lds_bytes = (int)lds;
  if (nb)
    VJ_HIP(hipMemcpy(P->d_order, order.data(), order.size() * sizeof(int), hipMemcpyHostToDevice));
  return 0;
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

// ev, if not null, holds NCLASS - 1 events recorded between the launches
// (autotuner timing)
template <typename T, bool CPLX>
int run_plan(const BPlan *P, void *Gr, void *Gi, void *Vr, void *Vi, void *eS, void *fS, vjh_int *info,
             uint64_t *transforms, hipStream_t s, hipEvent_t *ev = nullptr, unsigned mask = 7u)
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
  a.xfer = P->d_xfer;
  a.swp = P->swp;
  a.ord = P->ord;
  a.vpreset = P->vpreset;
  // a fill kernel, not hipMemsetAsync, so that the run can be graph-captured
  hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(1), 0, s, P->d_queue);
  // the long columns first (the costly matrices start first), then the
  // short ones on global memory, then the LDS-resident ones, each launch
  // with its own queue
  int first = 0;
  for (int c = 0; c < NCLASS; ++c) {
    a.nb = P->count[c];
    a.order = P->d_order + first;
    a.queue = P->d_queue + c;
    a.lds_bytes = (c == 2) ? P->lds_bytes : 0;
    a.group_lg = P->lg_c[c];
    if (mask & (1u << c))
      if (int rc = batched_svd<T, CPLX>(a, P->waves_c[c], P->max_m_c[c], 0, c != 0, s))
        return rc;
    first += P->count[c];
    if (ev && (c + 1 < NCLASS) && hipEventRecord(ev[c], s) != hipSuccess)
      return set_error(VJ_EHIP, "event record failed");
  }
  return 0;
}

int run_any(const BPlan *P, void *Gr, void *Gi, void *Vr, void *Vi, void *eS, void *fS, vjh_int *info,
            uint64_t *transforms, hipStream_t s, hipEvent_t *ev, unsigned mask = 7u)
{
  switch (P->kind) {
  case 0: return run_plan<double, false>(P, Gr, nullptr, Vr, nullptr, eS, fS, info, transforms, s, ev, mask);
  case 1: return run_plan<double, true>(P, Gr, Gi, Vr, Vi, eS, fS, info, transforms, s, ev, mask);
  case 2: return run_plan<float, false>(P, Gr, nullptr, Vr, nullptr, eS, fS, info, transforms, s, ev, mask);
  default: return run_plan<float, true>(P, Gr, Gi, Vr, Vi, eS, fS, info, transforms, s, ev, mask);
  }
}

// Launch shapes the autotuner tries: (lanes per pair as log2, waves per
// workgroup).  The wave form (class 0) has only the waves; the lane-group
// forms also take 8 and 4 lanes per pair where those are built.
struct Shape {
  int lg, waves;
};
std::vector<Shape> shape_candidates(const BPlan *P, int c)
{
  static const Shape all[9] = {{4, 1}, {4, 2}, {4, 4}, {4, 8}, {3, 1}, {3, 2}, {3, 4}, {2, 1}, {2, 2}};
  const bool cplx = (P->kind & 1) != 0;
  std::vector<Shape> v;
  for (const Shape &sh : all) {
    if (c == 0 && sh.lg != 4)
      continue;
    if (c != 0 && !group_shape_ok(cplx, P->max_m_c[c], sh.lg, sh.waves))
      continue;
    v.push_back(sh);
  }
  return v;
}

// Times the plan's launches (those in mask) at each launch shape on a
// scratch copy of G and keeps each launch's fastest; best[c] receives the
// time of that shape (ms, -1 for an empty class), and tot[k], if not null,
// the summed time of the launches at candidate k (k < 4: the waves 1, 2, 4,
// 8 at 16 lanes per pair).
int tune_shapes(BPlan *P, const void *G, const void *Gi, void *Gw, void *Gwi, void *V, void *Vi, void *eS, void *fS,
                vjh_int *info, unsigned mask, hipStream_t s, float *best, float *tot)
{
  const bool cplx = (P->kind & 1) != 0;
  const size_t bytes = (size_t)P->totG * (size_t)tsize(P->kind);
  hipEvent_t ev[NCLASS + 1] = {};
  int rc = 0;
  for (int i = 0; i <= NCLASS && !rc; ++i)
    if (hipEventCreate(&ev[i]) != hipSuccess)
      rc = set_error(VJ_EHIP, "autotune: event creation failed");
  std::vector<Shape> cand[NCLASS];
  size_t kmax = 0;
  Shape bs[NCLASS];
  for (int c = 0; c < NCLASS; ++c) {
    cand[c] = shape_candidates(P, c);
    kmax = std::max(kmax, cand[c].size());
    bs[c] = Shape{P->lg_c[c], P->waves_c[c]};
    best[c] = -1.0f;
  }
  for (size_t k = 0; k < kmax && !rc; ++k) {
    unsigned run_mask = 0u;
    for (int c = 0; c < NCLASS; ++c) {
      if (!(mask & (1u << c)) || !P->count[c] || k >= cand[c].size())
        continue;
      const int w = cap_waves(P->kind, P->max_m_c[c], cand[c][k].waves);
      if (w != cand[c][k].waves) // capped: the same launch as a smaller count
        continue;
      P->waves_c[c] = w;
      P->lg_c[c] = cand[c][k].lg;
      run_mask |= 1u << c;
    }
    if (tot && k < 4)
      tot[k] = -1.0f;
    if (!run_mask)
      continue;
    if (bytes && hipMemcpyAsync(Gw, G, bytes, hipMemcpyDeviceToDevice, s) != hipSuccess)
      rc = set_error(VJ_EHIP, "autotune: copy failed");
    if (!rc && bytes && cplx && hipMemcpyAsync(Gwi, Gi, bytes, hipMemcpyDeviceToDevice, s) != hipSuccess)
      rc = set_error(VJ_EHIP, "autotune: copy failed");
    if (!rc && hipEventRecord(ev[0], s) != hipSuccess)
      rc = set_error(VJ_EHIP, "autotune: event failed");
    if (!rc)
      rc = run_any(P, Gw, Gwi, V, Vi, eS, fS, info, nullptr, s, ev + 1, run_mask);
    if (!rc && (hipEventRecord(ev[NCLASS], s) != hipSuccess || hipEventSynchronize(ev[NCLASS]) != hipSuccess))
      rc = set_error(VJ_EHIP, "autotune: event failed");
    float sum = 0.0f;
    for (int c = 0; c < NCLASS && !rc; ++c) {
      float t = 0.0f;
      if (hipEventElapsedTime(&t, ev[c], ev[c + 1]) != hipSuccess) {
        rc = set_error(VJ_EHIP, "autotune: timing failed");
        break;
      }
      if (!(run_mask & (1u << c)))
        continue;
      sum += t;
      if (best[c] < 0.0f || t < best[c]) {
        best[c] = t;
        bs[c] = Shape{P->lg_c[c], P->waves_c[c]};
      }
    }
    if (tot && k < 4 && !rc)
      tot[k] = sum;
  }
  for (int c = 0; c < NCLASS; ++c) {
    P->lg_c[c] = bs[c].lg;
    P->waves_c[c] = bs[c].waves;
  }
  for (int i = 0; i <= NCLASS; ++i)
    if (ev[i])
      (void)hipEventDestroy(ev[i]);
  return rc;
}

// One timed run of the whole plan (every launch) on a scratch copy of G.
int time_plan(BPlan *P, const void *G, const void *Gi, void *Gw, void *Gwi, void *V, void *Vi, void *eS, void *fS,
              vjh_int *info, hipStream_t s, float *ms)
{
  const bool cplx = (P->kind & 1) != 0;
  const size_t bytes = (size_t)P->totG * (size_t)tsize(P->kind);
  hipEvent_t e0 = nullptr, e1 = nullptr;
  int rc = 0;
  if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess)
    rc = set_error(VJ_EHIP, "autotune: event creation failed");
  if (!rc && bytes && hipMemcpyAsync(Gw, G, bytes, hipMemcpyDeviceToDevice, s) != hipSuccess)
    rc = set_error(VJ_EHIP, "autotune: copy failed");
  if (!rc && bytes && cplx && hipMemcpyAsync(Gwi, Gi, bytes, hipMemcpyDeviceToDevice, s) != hipSuccess)
    rc = set_error(VJ_EHIP, "autotune: copy failed");
  if (!rc && hipEventRecord(e0, s) != hipSuccess)
    rc = set_error(VJ_EHIP, "autotune: event failed");
  if (!rc)
    rc = run_any(P, Gw, Gwi, V, Vi, eS, fS, info, nullptr, s, nullptr, 7u);
  if (!rc && (hipEventRecord(e1, s) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
              hipEventElapsedTime(ms, e0, e1) != hipSuccess))
    rc = set_error(VJ_EHIP, "autotune: event failed");
  if (e0)
    (void)hipEventDestroy(e0);
  if (e1)
    (void)hipEventDestroy(e1);
  return rc;
}

} // namespace

extern "C" {

void *vjh_bplan_create(int kind, vjh_int nb, const vjh_int *m, const vjh_int *n, const vjh_int *offG,
                       const vjh_int *offV, const vjh_int *offS, unsigned swp, int waves, vjh_int *status)
{
  vjh_int dummy = 0;
  vjh_int *st = status ? status : &dummy;
  *st = 0;
  if (kind < 0 || kind > 3 || nb < 0) {
    *st = set_error(-1, "bad kind or batch size");
    return nullptr;
  }
  std::vector<long long> hm(m, m + nb), hn(n, n + nb);
  int max_m = 0;
  for (vjh_int b = 0; b < nb; ++b) {
    if (hn[b] < 2 || (hn[b] & 1) || hn[b] > 128 || hm[b] < hn[b] || hm[b] > 1024 || offG[b] < 0 || offV[b] < 0 ||
        offS[b] < 0) {
      *st = set_error(-2, "matrix %lld: need even 2 <= n <= 128, n <= m <= 1024 and non-negative offsets (m=%lld n=%lld)",
                      (long long)b, hm[b], hn[b]);
      return nullptr;
    }
    max_m = std::max(max_m, (int)hm[b]);
  }
  long long totG = 0;
  for (vjh_int b = 0; b < nb; ++b)
    totG = std::max(totG, (long long)offG[b] + hm[b] * hn[b]);
  BPlan *P = new BPlan();
  P->kind = kind;
  P->nb = (int)nb;
  P->max_m = max_m;
  P->totG = totG;
  P->swp = swp;
  P->hm = hm;
  P->hn = hn;
  P->lds_budget = DEFAULT_LDS_BUDGET;
  P->split_rows = short_rows((kind & 1) != 0);
  const int w0 = (waves > 0) ? waves : (nb ? pick_waves(hn) : 4);
  (void)hipGetDevice(&P->dev);
  std::vector<long long> meta;
  meta.reserve((size_t)(5 * nb));
  meta.insert(meta.end(), hm.begin(), hm.end());
  meta.insert(meta.end(), hn.begin(), hn.end());
  meta.insert(meta.end(), offG, offG + nb);
  meta.insert(meta.end(), offV, offV + nb);
  meta.insert(meta.end(), offS, offS + nb);
  hipError_t e = hipSuccess;
  if (e == hipSuccess) e = hipMalloc((void **)&P->d_meta, std::max<size_t>(1, meta.size()) * sizeof(long long));
  if (e == hipSuccess) e = hipMalloc((void **)&P->d_order, std::max<size_t>(1, (size_t)nb) * sizeof(int));
  if (e == hipSuccess) e = hipMalloc((void **)&P->d_queue, NCLASS * sizeof(unsigned));
  if (e == hipSuccess) e = hipMalloc((void **)&P->d_evals, std::max<size_t>(1, (size_t)nb) * sizeof(unsigned long long));
  if (e == hipSuccess) e = hipMalloc((void **)&P->d_xfer, std::max<size_t>(1, (size_t)nb) * sizeof(unsigned long long));
  if (e == hipSuccess && !meta.empty()) e = hipMemcpy(P->d_meta, meta.data(), meta.size() * sizeof(long long), hipMemcpyHostToDevice);
  int rc = (e == hipSuccess) ? partition(P) : set_error(VJ_EHIP - (int)e, "plan allocation failed: %s", hipGetErrorString(e));
  if (rc) {
    *st = rc;
    (void)hipFree(P->d_meta);
    (void)hipFree(P->d_order);
    (void)hipFree(P->d_queue);
    (void)hipFree(P->d_evals);
    (void)hipFree(P->d_xfer);
    delete P;
    return nullptr;
  }
  for (int c = 0; c < NCLASS; ++c)
    P->waves_c[c] = cap_waves(kind, P->max_m_c[c], w0);
  return P;
}

vjh_int vjh_bplan_run(void *plan, void *Gr, void *Gi, void *Vr, void *Vi, void *eS, void *fS, vjh_int *info,
                      uint64_t *transforms, void *stream)
{
  const BPlan *P = (const BPlan *)plan;
  if (!P)
    return set_error(-1, "null plan");
  return run_any(P, Gr, Gi, Vr, Vi, eS, fS, info, transforms, (hipStream_t)stream, nullptr);
}

int vjh_bplan_waves(const void *plan)
{
  if (!plan)
    return 0;
  const BPlan *P = (const BPlan *)plan;
  for (int c = 0; c < NCLASS; ++c)
    if (P->count[c])
      return P->waves_c[c];
  return P->waves_c[0];
}

vjh_int vjh_bplan_config(const void *plan, vjh_int *out)
{
  const BPlan *P = (const BPlan *)plan;
  if (!P || !out)
    return set_error(-1, "null plan or output");
  for (int c = 0; c < NCLASS; ++c) {
    out[c] = P->count[c];
    out[NCLASS + c] = P->waves_c[c];
  }
  for (int c = 0; c < NCLASS; ++c)
    out[2 * NCLASS + c] = (c == 0) ? 64 : (1 << P->lg_c[c]);
  out[3 * NCLASS] = P->split_rows;
  out[3 * NCLASS + 1] = P->lds_budget;
  out[3 * NCLASS + 2] = P->lds_bytes;
  return 0;
}

vjh_int vjh_bplan_set(void *plan, int split_rows, vjh_int lds_budget, const int *waves, const int *lanes)
{
  BPlan *P = (BPlan *)plan;
  if (!P)
    return set_error(-1, "null plan");
  if (split_rows < 0 || split_rows > short_rows((P->kind & 1) != 0))
    return set_error(-2, "split row count %d outside 0..%d", split_rows, short_rows((P->kind & 1) != 0));
  if (lds_budget < 0 || lds_budget > LDS_FORM_MAX_BYTES)
    return set_error(-3, "LDS budget %lld outside 0..%d", (long long)lds_budget, LDS_FORM_MAX_BYTES);
  for (int c = 0; waves && c < NCLASS; ++c)
    if (waves[c] != 1 && waves[c] != 2 && waves[c] != 4 && waves[c] != 8)
      return set_error(-4, "waves per workgroup must be 1, 2, 4 or 8");
  for (int c = 1; lanes && c < NCLASS; ++c)
    if (lanes[c] != 16 && lanes[c] != 8 && lanes[c] != 4)
      return set_error(-5, "lanes per pair must be 16, 8 or 4");
  P->split_rows = split_rows;
  P->lds_budget = lds_budget;
  if (int rc = partition(P))
    return rc;
  for (int c = 0; waves && c < NCLASS; ++c)
    P->waves_c[c] = cap_waves(P->kind, P->max_m_c[c], waves[c]);
  for (int c = 1; lanes && c < NCLASS; ++c)
    P->lg_c[c] = (lanes[c] == 16) ? 4 : ((lanes[c] == 8) ? 3 : 2);
  const bool cplx = (P->kind & 1) != 0;
  for (int c = 1; c < NCLASS; ++c)
    if (P->count[c] && !group_shape_ok(cplx, P->max_m_c[c], P->lg_c[c], P->waves_c[c]))
      return set_error(-5, "class %d: %d lanes per pair with %d waves is not built for %d rows", c, 1 << P->lg_c[c],
                       P->waves_c[c], P->max_m_c[c]);
  return 0;
}

vjh_int vjh_bplan_evals(const void *plan, uint64_t *evals, void *stream)
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

vjh_int vjh_bplan_traffic(const void *plan, uint64_t *elems, void *stream)
{
  const BPlan *P = (const BPlan *)plan;
  if (!P || !elems)
    return set_error(-1, "null plan or output");
  hipStream_t s = (hipStream_t)stream;
  if (P->nb)
    VJ_HIP(hipMemcpyAsync(elems, P->d_xfer, (size_t)P->nb * sizeof(uint64_t), hipMemcpyDeviceToHost, s));
  VJ_HIP(hipStreamSynchronize(s));
  return 0;
}

vjh_int vjh_bplan_queue(void *plan, int model)
{
  BPlan *P = (BPlan *)plan;
  if (!P)
    return set_error(-1, "null plan");
  if (model < 0)
    return P->queue;
  if (model > 2)
    return set_error(-2, "batched SVD: queue cost model %d (0, 1 or 2)", model);
  P->queue = model;
  if (int rc = partition(P))
    return rc;
  return model;
}

vjh_int vjh_bplan_vpreset(void *plan, int on)
{
  BPlan *P = (BPlan *)plan;
  if (!P)
    return set_error(-1, "null plan");
  if (on >= 0)
    P->vpreset = on ? 1 : 0;
  return P->vpreset;
}

vjh_int vjh_bplan_order(void *plan, int id)
{
  BPlan *P = (BPlan *)plan;
  if (!P)
    return set_error(-1, "null plan");
  if (id == 0)
    return P->ord;
  if (id != 4 && id != 6)
    return set_error(-2, "batched SVD: ordering %d (pjs id 4 or 6)", id);
  P->ord = id;
  return id;
}

vjh_int vjh_bplan_autotune(void *plan, const void *G, const void *Gi, void *Gw, void *Gwi, void *V, void *Vi, void *eS,
                           void *fS, vjh_int *info, float *ms, void *stream)
{
  BPlan *P = (BPlan *)plan;
  if (!P)
    return set_error(-1, "null plan");
  hipStream_t s = (hipStream_t)stream;
  float best[NCLASS], tot[4];
  // 1. the row count above which a matrix takes the wave form: the default
  //    and half of it, every launch at its fastest waves per workgroup
  const int sr = short_rows((P->kind & 1) != 0);
  const int split_cand[2] = {sr, sr / 2};
  int best_split = P->split_rows;
  float best_split_ms = -1.0f;
  int best_split_waves[NCLASS], best_split_lg[NCLASS];
  for (int i = 0; i < 2; ++i) {
    P->split_rows = split_cand[i];
    if (int rc = partition(P))
      return rc;
    if (int rc = tune_shapes(P, G, Gi, Gw, Gwi, V, Vi, eS, fS, info, 7u, s, best, i == 0 ? tot : nullptr))
      return rc;
    float t = 0.0f;
    for (int c = 0; c < NCLASS; ++c)
      t += (best[c] > 0.0f) ? best[c] : 0.0f;
    if (ms && i == 0)
      for (int k = 0; k < 4; ++k)
        ms[k] = tot[k];
    if (best_split_ms < 0.0f || t < best_split_ms) {
      best_split_ms = t;
      best_split = split_cand[i];
      for (int c = 0; c < NCLASS; ++c) {
        best_split_waves[c] = P->waves_c[c];
        best_split_lg[c] = P->lg_c[c];
      }
    }
  }
  P->split_rows = best_split;
  for (int c = 0; c < NCLASS; ++c) {
    P->waves_c[c] = best_split_waves[c];
    P->lg_c[c] = best_split_lg[c];
  }
  // 2. the LDS budget: how many short matrices stay resident in LDS (the
  //    two short launches re-timed at each budget)
  static const long long budgets[5] = {0, 24 * 1024, 48 * 1024, 96 * 1024, LDS_FORM_MAX_BYTES};
  long long best_budget = P->lds_budget;
  float best_budget_ms = -1.0f;
  int bw1 = P->waves_c[1], bw2 = P->waves_c[2], bl1 = P->lg_c[1], bl2 = P->lg_c[2];
  const int w0 = P->waves_c[0];
  for (int i = 0; i < 5; ++i) {
    P->lds_budget = budgets[i];
    if (int rc = partition(P))
      return rc;
    if (int rc = tune_shapes(P, G, Gi, Gw, Gwi, V, Vi, eS, fS, info, 6u, s, best, nullptr))
      return rc;
    const float t = ((best[1] > 0.0f) ? best[1] : 0.0f) + ((best[2] > 0.0f) ? best[2] : 0.0f);
    if (best_budget_ms < 0.0f || t < best_budget_ms) {
      best_budget_ms = t;
      best_budget = budgets[i];
      bw1 = P->waves_c[1];
      bw2 = P->waves_c[2];
      bl1 = P->lg_c[1];
      bl2 = P->lg_c[2];
    }
  }
  P->lds_budget = best_budget;
  if (int rc = partition(P))
    return rc;
  P->waves_c[0] = w0;
  P->waves_c[1] = bw1;
  P->waves_c[2] = bw2;
  P->lg_c[1] = bl1;
  P->lg_c[2] = bl2;
  // 3. the processing order: each queue cost model, the whole plan timed
  //    twice at the chosen shapes, the fastest kept
  int best_q = P->queue;
  float best_q_ms = -1.0f;
  for (int q = 0; q < 3; ++q) {
    P->queue = q;
    if (int rc = partition(P))
      return rc;
    for (int rep = 0; rep < 2; ++rep) {
      float t = 0.0f;
      if (int rc = time_plan(P, G, Gi, Gw, Gwi, V, Vi, eS, fS, info, s, &t))
        return rc;
      if (best_q_ms < 0.0f || t < best_q_ms) {
        best_q_ms = t;
        best_q = q;
      }
    }
  }
  P->queue = best_q;
  if (int rc = partition(P))
    return rc;
  return (vjh_int)vjh_bplan_waves(P);
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
  (void)hipFree(P->d_xfer);
  delete P;
}

vjh_int vjh_dbvjsvd(vjh_int nb, const vjh_int *m, const vjh_int *n, const vjh_int *offG, const vjh_int *offV,
                    const vjh_int *offS, double *G, double *V, double *eS, double *fS, vjh_int *info,
                    uint64_t *transforms, unsigned swp, int waves, void *stream)
{
  vjh_int st = 0;
  void *P = vjh_bplan_create(0, nb, m, n, offG, offV, offS, swp, waves, &st);
  if (!P)
    return st;
  const vjh_int rc = vjh_bplan_run(P, G, nullptr, V, nullptr, eS, fS, info, transforms, stream);
  (void)hipStreamSynchronize((hipStream_t)stream);
  vjh_bplan_destroy(P);
  return rc;
}

} // extern "C"
