// Plain (one matrix) one-sided Jacobi SVD on gfx950: the device side of
// dvjsvd_/zvjsvd_/svjsvd_/cvjsvd_ (dvjsvd.c:40-638, zvjsvd.c:40-694).
//
// One kernel launch per step of the strategy table: workgroup k owns pair
// (js[st][2k], js[st][2k+1]) and runs the whole step on it (norms, dot,
// convergence test, 2x2 EVD, update of the G and V columns).  The control
// state (max-|.| bounds, power-of-two shifts, per-sweep transformation
// counter, convergence flag, error code) lives in device memory.  Each
// workgroup reports its pair in a slot; every workgroup of the next launch
// reduces the step's slots while its own pair loads and folds them into the
// control state (workgroup 0 stores the result), and a one-workgroup launch
// folds the last step of each sweep.  Sweeps run without any host round
// trip; once the convergence flag is set, later launches return at once.
#include "svd_common.hpp"

#include <algorithm>

namespace vj {

// ------------------------------------------------------------ setup kernels
template <typename T>
__global__ __launch_bounds__(256) void k_maxabs(const int m, const int n, const T *__restrict__ A, const long long ldA,
                                                unsigned long long *__restrict__ out)
{
  // max |a_ij| with NaN counted as inf (dznrmx.h:10-27)
  T v = (T)0;
  const long long total = (long long)m * n;
  for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < total; t += (long long)gridDim.x * blockDim.x) {
    const long long j = t / m, i = t - j * m;
    v = xmax(xmin(vabs(A[j * ldA + i]), inf_<T>()), v);
  }
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1)
    v = xmax(v, __shfl_xor(v, o, 64));
  if ((threadIdx.x & 63) == 0)
    atomicMax(out, (unsigned long long)__double_as_longlong((double)v));
}

template <typename T, bool CPLX>
__global__ void k_setup(const int m, PlainState<T> *st, const int vpreset)
{
  // dvjsvd.c:78-156 (zvjsvd.c:90-171): input checks and initial scaling
  if (threadIdx.x || blockIdx.x)
    return;
  const int L_MG = CPLX ? 92 : 80, L_MV = CPLX ? 104 : 92;
  const T MG = (T)__longlong_as_double((long long)st->MG_u);
  st->MG0 = MG;
  st->nev = 0u;
  st->err_sw = st->err_st = 0u;
  const T MV = vpreset ? (T)__longlong_as_double((long long)st->MV_u) : (T)1;
  st->sT = st->sV = 0;
  st->ctl[0].sG = st->ctl[0].sV = 0;
  st->sG_init = st->sV_init = 0;
  st->sweeps = 0;
  st->vsweeps = 0;
  st->swt = 0;
  st->zero = 0;
  st->status = 0;
  st->done = 0;
  st->ef_init = 0;
  if (!(MG <= Lim<T>::MAX)) { st->status = -L_MG; st->done = 1; return; }
  if (vpreset) {
    if (!(MV <= Lim<T>::MAX)) { st->status = -L_MV; st->done = 1; return; }
    if (MV <= (T)0) { st->status = CPLX ? -17 : -13; st->done = 1; return; }
  }
  st->ef_init = 1; // V (unless preset) and eS/fS are initialised from here on
  if (MG == (T)0) { st->zero = 1; st->done = 1; return; }
  const T M_m = CPLX ? (Lim<T>::MAX / ((T)(m << 2) * (T)1.41421356237309504880)) : (Lim<T>::MAX / (T)(m << 1));
  const int NRM_EXP = frexp_exp(M_m) - 1;
  const int eM = frexp_exp(MG) - 1;
  const int sN = NRM_EXP - eM - 1;
  st->eM0 = eM;
  st->sR0 = (CPLX ? Lim<T>::ROT_EXP_C - 1 : Lim<T>::ROT_EXP_R) - eM;
  st->NRM_EXP = NRM_EXP;
  st->ctl[0].MG = ldexp_(MG, sN);
  st->sG_init = sN;
  st->sT = sN;
  int sV = 0;
  if (vpreset) {
    const int eV = frexp_exp(MV) - 1;
    sV = (CPLX ? Lim<T>::ROT_EXP_C - 1 : Lim<T>::ROT_EXP_R) - eV;
    st->ctl[0].MV = ldexp_(MV, sV);
  }
  else
    st->ctl[0].MV = (T)1;
  st->sV_init = sV;
  st->sV = sV;
}

template <typename T>
__global__ __launch_bounds__(256) void k_init_cols(const int m, const int n, T *__restrict__ Gr, T *__restrict__ Gi,
                                                   const long long ldG, T *__restrict__ Vr, T *__restrict__ Vi,
                                                   const long long ldV, T *__restrict__ eS, T *__restrict__ fS,
                                                   const PlainState<T> *__restrict__ st, const int vpreset)
{
  // column j = blockIdx.x: initial power-of-two scaling of G (dscale_), of a
  // preset V, V := I otherwise (dvjsvd.c:97-156), eS = -inf, fS = 1
  if (!st->ef_init)
    return;
  const int j = blockIdx.x;
  const int sG = st->zero ? 0 : st->sG_init;
  const int sV = st->sV_init;
  if (sG) {
    for (int i = threadIdx.x; i < m; i += blockDim.x) {
      Gr[j * ldG + i] = ldexp_(Gr[j * ldG + i], sG);
      if (Gi)
        Gi[j * ldG + i] = ldexp_(Gi[j * ldG + i], sG);
    }
  }
  for (int i = threadIdx.x; i < n; i += blockDim.x) {
    if (vpreset) {
      if (sV) {
        Vr[j * ldV + i] = ldexp_(Vr[j * ldV + i], sV);
        if (Vi)
          Vi[j * ldV + i] = ldexp_(Vi[j * ldV + i], sV);
      }
    }
    else {
      Vr[j * ldV + i] = (i == j) ? (T)1 : (T)0;
      if (Vi)
        Vi[j * ldV + i] = (T)0;
    }
  }
  if (threadIdx.x == 0) {
    eS[j] = -inf_<T>();
    fS[j] = (T)1;
  }
}

// -------------------------------------------------- deferred V updates
// With V := I (no preset V, so no tracking or rescaling of V) the steps
// leave V alone and record each pair's decision; after every sweep k_vflush
// applies the sweep's steps to V in order, R rows at a time held in LDS.  An
// element of V only meets the element of the same row in the other column of
// its pair, so each goes through the same operations in the same order as
// in the steps: the same bits, for one pass over V per sweep instead of a
// read and a write of both V columns of every transformed pair at every step.
template <typename T> struct VRec {
  int *code; // action code per step and pair (nullptr: V updated in the steps)
  T *c, *at, *sat;
};

// Workgroup barrier that orders the LDS accesses only (global loads in
// flight stay in flight).
__device__ __forceinline__ void lds_barrier()
{
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
  __builtin_amdgcn_s_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// One element pair of V under a recorded decision, as pair_apply updates
// the V columns (codes 2 and -2: rotation, with a zero tangent a swap when
// the cosine is 1 and the pair is permuted; -1 and -3: swap; 1 and 3: none).
template <typename T, bool CPLX>
__device__ __forceinline__ void vflush_elem(const int code, const T c, const T at, const T sat, const int sec, T &xr,
                                            T &yr, T &xi, T &yi)
{
  bool swap = (code == -1) || (code == -3);
  if ((code == 2) || (code == -2)) {
    const bool perm = (code < 0);
    if ((at == (T)0) && (!CPLX || (sat == (T)0)))
      swap = (c == (T)1) && perm;
    else {
      const bool c1 = (c == (T)1);
      const T x_r = xr, y_r = yr;
      if (!CPLX || (sat == (T)0)) {
        T a = perm ? fma_(-x_r, at, y_r) : fma_(y_r, at, x_r);
        T b = perm ? fma_(y_r, at, x_r) : fma_(-x_r, at, y_r);
        if (!c1) { a = cscale(a, c, sec); b = cscale(b, c, sec); }
        xr = a;
        yr = b;
        if constexpr (CPLX) {
          const T x_i = xi, y_i = yi;
          T a2 = perm ? fma_(-x_i, at, y_i) : fma_(y_i, at, x_i);
          T b2 = perm ? fma_(y_i, at, x_i) : fma_(-x_i, at, y_i);
          if (!c1) { a2 = cscale(a2, c, sec); b2 = cscale(b2, c, sec); }
          xi = a2;
          yi = b2;
        }
      }
      else if constexpr (CPLX) {
        const T cna = -at;
        const T x_i = xi, y_i = yi;
        T ar = fma_(y_r, at, fma_(-y_i, sat, x_r));
        T ai = fma_(y_r, sat, fma_(y_i, at, x_i));
        T br = fma_(x_r, cna, fma_(-x_i, sat, y_r));
        T bi = fma_(x_r, sat, fma_(x_i, cna, y_i));
        if (!c1) { ar = cscale(ar, c, sec); ai = cscale(ai, c, sec); br = cscale(br, c, sec); bi = cscale(bi, c, sec); }
        if (perm) { xr = br; xi = bi; yr = ar; yi = ai; }
        else { xr = ar; xi = ai; yr = br; yi = bi; }
      }
    }
  }
  if (swap) {
    T t = xr; xr = yr; yr = t;
    if constexpr (CPLX) { t = xi; xi = yi; yi = t; }
  }
}

// A pair's recorded decision for one step, with its column indices.
template <typename T> struct VFDec {
  bool ok; // the pair and the step exist
  int code;
  unsigned p, q;
  T c, at, sat;
};

constexpr int VF_LDS_MAX = 64 * 1024; // LDS per workgroup (two per CU)
constexpr int VF_DEPTH = 2;           // steps whose decisions are loaded ahead

// Rows [R blockIdx.x, R blockIdx.x + R) of V through the steps of sweep sw
// (when its steps ran: st->vsweeps == sw + 1; after a failing step only the
// steps up to it).  LDS holds the rows column by column (element (r, j) at
// j R + r); the lanes take the pairs of a step (PM each), so a step's
// element pairs are independent and one LDS barrier separates the steps.
// The decisions of the next step are loaded while a step is applied.
template <typename T, bool CPLX, int PM>
__global__ __launch_bounds__(1024) void k_vflush(const int nv, const int n, const int R, T *__restrict__ Vr,
                                                 T *__restrict__ Vi, const long long ldV,
                                                 const unsigned *__restrict__ js, const unsigned stp, const VRec<T> rec,
                                                 const PlainState<T> *__restrict__ st, const unsigned sw, const int sec)
{
  extern __shared__ double vf_lds[];
  if (st->vsweeps != sw + 1u)
    return;
  const unsigned nst = (st->status && (st->err_sw == sw)) ? st->err_st + 1u : stp;
  if (nst == 0u)
    return;
  T *const br = (T *)vf_lds;
  T *const bi = br + (size_t)R * n;
  const int i0 = blockIdx.x * R;
  const int rows = (nv - i0 < R) ? (nv - i0) : R;
  const int lane = threadIdx.x, nt = blockDim.x;
  for (int t = lane; t < R * n; t += nt) {
    const int r = t % R, j = t / R;
    const bool in = (r < rows);
    br[t] = in ? Vr[(long long)j * ldV + i0 + r] : (T)0;
    if constexpr (CPLX)
      bi[t] = in ? Vi[(long long)j * ldV + i0 + r] : (T)0;
  }
  const int np = n >> 1;
  // VF_DEPTH sets of decisions in registers: a step is applied while the
  // loads of the next VF_DEPTH - 1 steps are in flight
  VFDec<T> d[VF_DEPTH][PM];
  auto fetch = [&](const unsigned s, VFDec<T>(&e)[PM]) {
#pragma unroll
    for (int i = 0; i < PM; ++i) {
      const int k = lane + i * nt;
      const bool ok = (k < np) && (s < nst);
      const long long ix = ok ? (long long)s * np + k : 0;
      const size_t jx = ok ? (size_t)s * n + 2 * k : 0;
      e[i].ok = ok;
      e[i].code = rec.code[ix];
      e[i].c = rec.c[ix];
      e[i].at = rec.at[ix];
      e[i].sat = CPLX ? rec.sat[ix] : (T)0;
      e[i].p = js[jx];
      e[i].q = js[jx + 1];
    }
  };
  auto apply = [&](const VFDec<T>(&e)[PM]) {
#pragma unroll
    for (int i = 0; i < PM; ++i) {
      const int code = e[i].code;
      if (!e[i].ok || (code == 1) || (code == 3))
        continue;
      T *const xp = br + (size_t)e[i].p * R, *const yp = br + (size_t)e[i].q * R;
      T *const xpi = bi + (size_t)e[i].p * R, *const ypi = bi + (size_t)e[i].q * R;
      for (int r = 0; r < R; ++r) {
        T xr = xp[r], yr = yp[r], xi = (T)0, yi = (T)0;
        if constexpr (CPLX) {
          xi = xpi[r];
          yi = ypi[r];
        }
        vflush_elem<T, CPLX>(code, e[i].c, e[i].at, e[i].sat, sec, xr, yr, xi, yi);
        xp[r] = xr;
        yp[r] = yr;
        if constexpr (CPLX) {
          xpi[r] = xi;
          ypi[r] = yi;
        }
      }
    }
  };
#pragma unroll
  for (int u = 0; u < VF_DEPTH; ++u) {
    fetch((unsigned)u, d[u]);
    __builtin_amdgcn_sched_barrier(0); // the loads issue in step order
  }
  __syncthreads();
  // (a step past the last one applies nothing: every lane runs the same
  // barriers, and the loads are issued in one order on every path)
  for (unsigned s = 0; s < nst; s += VF_DEPTH) {
#pragma unroll
    for (int u = 0; u < VF_DEPTH; ++u) {
      apply(d[u]);
      lds_barrier();
      fetch(s + VF_DEPTH + u, d[u]);
    }
  }
  for (int t = lane; t < R * n; t += nt) {
    const int r = t % R, j = t / R;
    if (r < rows) {
      Vr[(long long)j * ldV + i0 + r] = br[t];
      if constexpr (CPLX)
        Vi[(long long)j * ldV + i0 + r] = bi[t];
    }
  }
}

// ------------------------------------------------------------- step kernel
// Rows lane + j B (j < E) of the pair's columns, zero past m, scaled by
// 2^sG when a rescaling is pending.
template <typename T, bool CPLX, int B, int E>
__device__ __forceinline__ void load_rows(const T *Gp, const T *Gq, const T *Gpi, const T *Gqi, const int m, const int j0,
                                          const int sG, const int lane, T (&xr)[E], T (&yr)[E], T (&xi)[E], T (&yi)[E])
{
#pragma unroll
  for (int j = 0; j < E; ++j) {
    const int i = lane + (j0 + j) * B;
    const bool in = (i < m);
    xr[j] = in ? Gp[i] : (T)0;
    yr[j] = in ? Gq[i] : (T)0;
    if constexpr (CPLX) {
      xi[j] = in ? Gpi[i] : (T)0;
      yi[j] = in ? Gqi[i] : (T)0;
    }
    else {
      xi[j] = yi[j] = (T)0;
    }
    if (sG) {
      xr[j] = ldexp_(xr[j], sG);
      yr[j] = ldexp_(yr[j], sG);
      if constexpr (CPLX) {
        xi[j] = ldexp_(xi[j], sG);
        yi[j] = ldexp_(yi[j], sG);
      }
    }
  }
}

// A store that a later launch reads, written through the L2 (an agent-scope
// relaxed store): the launch then ends with fewer dirty lines to write back
// before the next one starts (on 1024 x 1024 FP64, 147.4 -> 137.9 ms per
// SVD; plain and non-temporal stores leave it at 147-148 ms).
template <typename T> __device__ __forceinline__ void st_wt(T *p, const T v)
{
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

template <typename T, bool CPLX, int B, int E>
__device__ __forceinline__ void store_rows(T *Gp, T *Gq, T *Gpi, T *Gqi, const int m, const int j0, const int lane,
                                           const T (&xr)[E], const T (&yr)[E], const T (&xi)[E], const T (&yi)[E])
{
#pragma unroll
  for (int j = 0; j < E; ++j) {
    const int i = lane + (j0 + j) * B;
    if (i < m) {
      st_wt(Gp + i, xr[j]);
      st_wt(Gq + i, yr[j]);
      if constexpr (CPLX) {
        st_wt(Gpi + i, xi[j]);
        st_wt(Gqi + i, yi[j]);
      }
    }
  }
}

// Power-of-two scaling of E chunks held in registers (a rescaling decided
// after the chunks were loaded).
template <typename T, bool CPLX, int E>
__device__ __forceinline__ void scale_rows(const int s, T (&xr)[E], T (&yr)[E], T (&xi)[E], T (&yi)[E])
{
#pragma unroll
  for (int j = 0; j < E; ++j) {
    xr[j] = ldexp_(xr[j], s);
    yr[j] = ldexp_(yr[j], s);
    if constexpr (CPLX) {
      xi[j] = ldexp_(xi[j], s);
      yi[j] = ldexp_(yi[j], s);
    }
  }
}

// A streamed step keeps the last Keep::N chunks its first pass reads in LDS
// (value j of a lane at [j][lane]), and its update pass, which visits them
// first, takes them from there instead of reading them again: 32 KiB per
// workgroup, within the LDS the streamed shapes' occupancy leaves free.
template <typename T, bool CPLX, int B, int E> struct Keep {
  static constexpr int V = E * (CPLX ? 4 : 2);                   // values per lane per chunk
  static constexpr int N = (32768 / (int)sizeof(T)) / (V * B);    // chunks kept
};

template <typename T, bool CPLX, int B, int E>
__device__ __forceinline__ void keep_rows(T *kb, const int lane, const T (&xr)[E], const T (&yr)[E], const T (&xi)[E],
                                          const T (&yi)[E])
{
  constexpr int NV = CPLX ? 4 : 2; // values per row
#pragma unroll
  for (int j = 0; j < E; ++j) {
    kb[(NV * j + 0) * B + lane] = xr[j];
    kb[(NV * j + 1) * B + lane] = yr[j];
    if constexpr (CPLX) {
      kb[(NV * j + 2) * B + lane] = xi[j];
      kb[(NV * j + 3) * B + lane] = yi[j];
    }
  }
}

template <typename T, bool CPLX, int B, int E>
__device__ __forceinline__ void kept_rows(const T *kb, const int lane, T (&xr)[E], T (&yr)[E], T (&xi)[E], T (&yi)[E])
{
  constexpr int NV = CPLX ? 4 : 2;
#pragma unroll
  for (int j = 0; j < E; ++j) {
    xr[j] = kb[(NV * j + 0) * B + lane];
    yr[j] = kb[(NV * j + 1) * B + lane];
    if constexpr (CPLX) {
      xi[j] = kb[(NV * j + 2) * B + lane];
      yi[j] = kb[(NV * j + 3) * B + lane];
    }
    else {
      xi[j] = yi[j] = (T)0;
    }
  }
}

// --------------------------------------------------------- fold of a step
// Maxima, transformation count and first failure over the slots of one
// step, returned to every lane (maxima, an integer sum and a minimum do not
// depend on the order of the reduction).
struct SlotTotals {
  double tG, tV, cosv, sinv;
  unsigned long long trans;
  unsigned long long key; // (pair index << 32 | code) of the first failing pair, ~0 if none
};

// The fields of SP slots per lane that every fold reads (error code,
// maxima, transformation flag), requested together at the start of a step:
// they do not depend on the step's pair, and no load waits for another (the
// error code does not gate the loads of the maxima).
template <int SP> struct SlotPre {
  int err[SP];
  unsigned trans[SP];
  double tG[SP], tV[SP];
};
template <> struct SlotPre<0> {};

template <int B, int SP>
__device__ __forceinline__ void prefetch_slots(const PlainSlot *__restrict__ sl, const int np, const int lane,
                                               SlotPre<SP> &pf)
{
#pragma unroll
  for (int u = 0; u < SP; ++u) {
    const int i = lane + u * B;
    const PlainSlot *const e = sl + ((i < np) ? i : 0);
    pf.err[u] = e->err;
    pf.trans[u] = e->trans;
    pf.tG[u] = e->tG;
    pf.tV[u] = e->tV;
  }
}

// Marks the prefetched slots as received at this point: their loads stay
// above it (they would otherwise be sunk to their first use, after the
// pair's columns are requested) and their use below it.
template <int SP> __device__ __forceinline__ void pin_slots(SlotPre<SP> &pf)
{
  if constexpr (SP > 0) {
#pragma unroll
    for (int u = 0; u < SP; ++u)
      asm volatile("" : "+v"(pf.err[u]), "+v"(pf.trans[u]), "+v"(pf.tG[u]), "+v"(pf.tV[u]));
  }
}

__device__ __forceinline__ void fold_slot(SlotTotals &t, const PlainSlot *__restrict__ sl, const int i, const int e,
                                          const double tG, const double tV, const unsigned trans, const int stats)
{
  if (e) {
    const unsigned long long key = ((unsigned long long)(unsigned)i << 32) | (unsigned)e;
    t.key = (key < t.key) ? key : t.key;
    return;
  }
  t.tG = fmax(t.tG, tG);
  t.tV = fmax(t.tV, tV);
  t.trans += trans;
  if (stats) {
    t.cosv = fmax(t.cosv, sl[i].cosv);
    t.sinv = fmax(t.sinv, sl[i].sinv);
  }
}

// Slots [0, np): the first SP B from pf, the rest loaded here.  POST =
// false leaves out the barrier after the lds reads (the caller uses other
// lds areas until its next barrier).
template <int B, int SP, bool POST = true>
__device__ __forceinline__ SlotTotals reduce_slots(const PlainSlot *__restrict__ sl, const int np, const int stats,
                                                   const int lane, double *lds, const SlotPre<SP> &pf)
{
  SlotTotals t{0.0, 0.0, 0.0, 0.0, 0ull, ~0ull};
  if constexpr (SP > 0) {
#pragma unroll
    for (int u = 0; u < SP; ++u) {
      const int i = lane + u * B;
      if (i < np)
        fold_slot(t, sl, i, pf.err[u], pf.tG[u], pf.tV[u], pf.trans[u], stats);
    }
  }
  for (int i = lane + SP * B; i < np; i += B) {
    const int e = sl[i].err;
    const unsigned tr = sl[i].trans;
    const double g = sl[i].tG, v = sl[i].tV;
    fold_slot(t, sl, i, e, g, v, tr, stats);
  }
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) {
    t.tG = fmax(t.tG, __shfl_xor(t.tG, o, 64));
    t.tV = fmax(t.tV, __shfl_xor(t.tV, o, 64));
    t.cosv = fmax(t.cosv, __shfl_xor(t.cosv, o, 64));
    t.sinv = fmax(t.sinv, __shfl_xor(t.sinv, o, 64));
    t.trans += __shfl_xor(t.trans, o, 64);
    const unsigned long long k2 = __shfl_xor(t.key, o, 64);
    t.key = (k2 < t.key) ? k2 : t.key;
  }
  if constexpr (B > 64) {
    double *const w = lds + 6 * (lane >> 6);
    if ((lane & 63) == 0) {
      w[0] = t.tG; w[1] = t.tV; w[2] = t.cosv; w[3] = t.sinv;
      w[4] = __longlong_as_double((long long)t.trans);
      w[5] = __longlong_as_double((long long)t.key);
    }
    __syncthreads();
    SlotTotals a{0.0, 0.0, 0.0, 0.0, 0ull, ~0ull};
    for (int w2 = 0; w2 < B / 64; ++w2) {
      const double *const u = lds + 6 * w2;
      a.tG = fmax(a.tG, u[0]); a.tV = fmax(a.tV, u[1]); a.cosv = fmax(a.cosv, u[2]); a.sinv = fmax(a.sinv, u[3]);
      a.trans += (unsigned long long)__double_as_longlong(u[4]);
      const unsigned long long k2 = (unsigned long long)__double_as_longlong(u[5]);
      a.key = (k2 < a.key) ? k2 : a.key;
    }
    if constexpr (POST)
      __syncthreads(); // lds is reused by the step
    t = a;
  }
  return t;
}

template <typename T> struct Fold {
  PlainCtl<T> c;  // the state before the next step
  int eM, sR;     // exponent of MG and the rotation headroom (trace)
  bool shiftG;    // a rescaling of G was decided
};

// dvjsvd.c:198-233: fold a step's maxima into the running bounds and decide
// the power-of-two rescaling of G and V for the next step.
template <typename T, bool CPLX>
__device__ __forceinline__ Fold<T> fold_ctl(const PlainCtl<T> &in, const SlotTotals &t, const int NRM_EXP)
{
  const T nMG = (T)t.tG, nMV = (T)t.tV;
  T MG = (nMG > in.MG) ? nMG : in.MG;
  T MV = (nMV > in.MV) ? nMV : in.MV;
  const int rot = CPLX ? Lim<T>::ROT_EXP_C - 1 : Lim<T>::ROT_EXP_R;
  Fold<T> f;
  f.eM = frexp_exp(MG) - 1;
  f.sR = rot - f.eM;
  f.shiftG = (f.sR < 0);
  int sG = 0, sV = 0;
  if (f.shiftG) {
    sG = NRM_EXP - f.eM - 1;
    MG = ldexp_(MG, sG);
  }
  const int eV = frexp_exp(MV) - 1;
  if (rot - eV < 0) {
    sV = rot - eV;
    MV = ldexp_(MV, sV);
  }
  f.c.MG = MG;
  f.c.MV = MV;
  f.c.sG = sG;
  f.c.sV = sV;
  return f;
}

// The stores of the fold of step `folded` (one lane of one workgroup): the
// next ctl, the shift totals and their trace, the sweep's transformation
// count and, after the last step of a sweep, the end-of-sweep test
// (dvjsvd.c:599-601).  A failing pair ends the run with its code.
template <typename T>
__device__ void fold_commit(PlainState<T> *st, const unsigned r, const SlotTotals &t, const Fold<T> &f,
                            const unsigned folded, const unsigned stp, const unsigned swp, const int stats)
{
  if (t.key != ~0ull) {
    st->status = (int)(unsigned)(t.key & 0xffffffffull);
    st->err_sw = st->sweeps;
    st->err_st = folded;
    st->ctl[r ^ 1u] = st->ctl[r];
    st->done = 1;
    return;
  }
  st->swt += t.trans;
  if (stats) {
    st->cos_max = fmax(st->cos_max, t.cosv);
    st->sin_max = fmax(st->sin_max, t.sinv);
  }
  const bool wrap = (folded + 1u == stp);
  if (f.shiftG) {
    st->sT += f.c.sG;
    if (st->nev < 32u) { // logged against the step it precedes
      auto &e = st->ev[st->nev++];
      e.sw = st->sweeps + (wrap ? 1u : 0u);
      e.st = wrap ? 0u : folded + 1u;
      e.eM = f.eM;
      e.sR = f.sR;
      e.sN = f.c.sG;
      e.M = f.c.MG;
    }
  }
  st->sV += f.c.sV;
  st->ctl[r ^ 1u] = f.c;
  if (wrap) {
    if (stats) {
      st->last_cos = st->cos_max;
      st->last_sin = st->sin_max;
      st->cos_max = st->sin_max = 0.0;
    }
    const unsigned long long swt = st->swt;
    if (st->sweeps < 64u)
      st->hist[st->sweeps] = swt;
    if (!swt)
      st->done = 1;
    else {
      st->sweeps += 1u;
      st->swt = 0ull;
      if (st->sweeps >= swp)
        st->done = 1;
    }
  }
}

// The fold of the last step of a sweep, in a launch of its own (one
// workgroup), so that the convergence flag is current when the sweep ends.
template <typename T, bool CPLX>
__global__ __launch_bounds__(256) void k_fold(PlainState<T> *__restrict__ st, const PlainSlot *__restrict__ prev,
                                              const int np, const unsigned r, const unsigned stp, const unsigned swp,
                                              const int stats)
{
  __shared__ double lds[6 * 4];
  const int lane = threadIdx.x;
  if (st->done) {
    if (lane == 0)
      st->ctl[r ^ 1u] = st->ctl[r];
    return;
  }
  const SlotTotals t = reduce_slots<256, 0>(prev, np, stats, lane, lds, SlotPre<0>{});
  if (lane == 0)
    fold_commit<T>(st, r, t, fold_ctl<T, CPLX>(st->ctl[r], t, st->NRM_EXP), stp - 1u, stp, swp, stats);
}

// One step on pair blockIdx.x.  STREAM = false: the pair is held in
// registers, E rows per lane.  STREAM = true (columns too long for the
// registers): the pair is read three times in chunks of E rows per lane -
// maxima, sums, update - with the same per-lane row order, hence the same
// bits as the register form with the same group size.
template <typename T, bool CPLX, int B, int E, bool STREAM = false>
__global__ __launch_bounds__(B) void k_step(const int m, const int n, T *__restrict__ Gr, T *__restrict__ Gi,
                                             const long long ldG, T *__restrict__ Vr, T *__restrict__ Vi,
                                             const long long ldV, T *__restrict__ eS, T *__restrict__ fS,
                                             const unsigned *__restrict__ row, const unsigned st_idx,
                                             const unsigned stp, const unsigned swp, const T tol, const T gst,
                                             const int vtrack, const int stats, const int sec,
                                             PlainState<T> *__restrict__ st,
                                             const unsigned r, const PlainSlot *__restrict__ prev,
                                             PlainSlot *__restrict__ slots, const VRec<T> vrec)
{
  // n is the length of the V columns (the table row has gridDim.x pairs);
  // n = 0 when the V updates are deferred (vrec.code set): the step then
  // records its decision for k_vflush instead.
  // The control state before this step is ctl[r] with, for st_idx > 0, the
  // slots of step st_idx - 1 (prev) folded in; workgroup 0 stores the fold
  // in ctl[r ^ 1].
  // lds areas of the group reductions: the fold of the slots (6 per wave),
  // then, in the register form, the maxima (2), the sums (8) and the update
  // maximum (1) each in an area of its own, so that none needs a barrier
  // before its writes
  constexpr int W64 = B / 64;
  constexpr int L_MAX2 = 6 * W64, L_DD4 = 8 * W64, L_TG = 16 * W64;
  __shared__ double lds[17 * W64 + 2];
  // the streamed form's kept chunks (each lane reads back only what it wrote)
  __shared__ T kbuf[(STREAM && (Keep<T, CPLX, B, E>::N > 0)) ? Keep<T, CPLX, B, E>::N * Keep<T, CPLX, B, E>::V * B : 1];
  const int lane = threadIdx.x;
  const bool fold = (st_idx > 0u);
  // The pair and the flag are read together (one latency, not two).  A plain
  // load of the flag is enough: it was set by an earlier launch, or, after a
  // failing pair, by workgroup 0 of this one, and then every workgroup finds
  // that failure in the slots itself.
  const int k = blockIdx.x;
  // the previous step's slots are requested first: they do not depend on
  // the pair, so their latency overlaps that of the pair's indices and then
  // of its columns
  // (unconditionally: the first step of a sweep reads the other half of the
  // slot buffer, allocated, whose contents it does not use, so that no
  // branch separates the loads from their use)
  constexpr int SP = (B >= 256) ? 2 : 4;
  const int np = (int)gridDim.x;
  // the arguments used before the pair's columns are requested, read in one batch
  asm volatile("" ::"s"(row), "s"(st), "s"(prev), "s"(np), "s"(r), "s"(m), "s"(Gr), "s"(ldG));
  if constexpr (CPLX)
    asm volatile("" ::"s"(Gi));
  SlotPre<SP> pf;
  prefetch_slots<B, SP>(prev, np, lane, pf);
  const int p = (int)row[2 * k], q = (int)row[2 * k + 1];
  const int done = st->done;
  PlainCtl<T> c = st->ctl[r];
  const int nrm = st->NRM_EXP;
  // the indices, the flag and the control state issued together before the
  // test below, and the slots received with them (requested at the same time)
  asm volatile("" ::"s"(p), "s"(q), "s"(done), "s"(c.MG), "s"(c.MV), "s"(c.sG), "s"(c.sV), "s"(nrm));
  pin_slots<SP>(pf);
  if (done) {
    if (fold && (blockIdx.x == 0) && (lane == 0))
      st->ctl[r ^ 1u] = c;
    return;
  }

  T *const Gp = Gr + (long long)p * ldG, *const Gq = Gr + (long long)q * ldG;
  T *const Gpi = CPLX ? Gi + (long long)p * ldG : nullptr, *const Gqi = CPLX ? Gi + (long long)q * ldG : nullptr;
  T *const Vp = Vr + (long long)p * ldV, *const Vq = Vr + (long long)q * ldV;
  T *const Vpi = CPLX ? Vi + (long long)p * ldV : nullptr, *const Vqi = CPLX ? Vi + (long long)q * ldV : nullptr;
  // the V pair in registers along with G (loaded at the start, so that its
  // latency overlaps the reductions) when it fits E chunks of B rows (E <= 8,
  // complex E <= 4)
  const bool vreg = !STREAM && (E * (CPLX ? 2 : 1) <= 8) && (n <= E * B);
  T xr[E], yr[E], xi[E], yi[E], vx[E], vy[E], vxi[E], vyi[E];
  // The pair (and the V pair) is requested before the previous step's slots
  // are reduced, so that the two latencies overlap; a pending rescaling is
  // applied once the fold has decided it.
  if (!STREAM)
    load_rows<T, CPLX, B, E>(Gp, Gq, Gpi, Gqi, m, 0, 0, lane, xr, yr, xi, yi);
  if (vreg)
    load_rows<T, CPLX, B, E>(Vp, Vq, Vpi, Vqi, n, 0, 0, lane, vx, vy, vxi, vyi);
  if (fold) {
    const SlotTotals t = reduce_slots<B, SP, false>(prev, np, stats, lane, lds, pf);
    const Fold<T> f = fold_ctl<T, CPLX>(c, t, nrm);
    if ((k == 0) && (lane == 0))
      fold_commit<T>(st, r, t, f, st_idx - 1u, stp, swp, stats);
    if (t.key != ~0ull)
      return;
    c = f.c;
  }
  Step<T> S;
  S.tol = tol;
  S.gst = gst;
  S.sG = c.sG;
  S.sV = c.sV;
  S.vtrack = vtrack;
  S.sec = sec;
  if (!STREAM && S.sG)
    scale_rows<T, CPLX, E>(S.sG, xr, yr, xi, yi);
  if (vreg && S.sV)
    scale_rows<T, CPLX, E>(S.sV, vx, vy, vxi, vyi);
  if (S.sV && !vreg) {
    for (int i = lane; i < n; i += B) {
      Vp[i] = ldexp_(Vp[i], S.sV);
      Vq[i] = ldexp_(Vq[i], S.sV);
      if constexpr (CPLX) {
        Vpi[i] = ldexp_(Vpi[i], S.sV);
        Vqi[i] = ldexp_(Vqi[i], S.sV);
      }
    }
  }
  T eSp, fSp, eSq, fSq;
  PairOut<T, CPLX> o;
  PairDec<T> D;
  if (vreg) {
    const int nEV = (n + B - 1) / B;
    Step<T> Sg = S;
    Sg.vtrack = 0;
    {
      // pair_step with its reductions in their own lds areas
      typedef typename Acc<T>::A A;
      T mx, my;
      pair_maxima<T, CPLX, E>(xr, yr, xi, yi, E, mx, my);
      red_max2<T, B, false>(mx, my, lds + L_MAX2);
      const int kx = max_exp(mx), ky = max_exp(my);
      DD<A> sx{0, 0}, sy{0, 0}, dr{0, 0}, di{0, 0};
      pair_sums<T, CPLX, E>(xr, yr, xi, yi, E, kx, ky, sx, sy, dr, di);
      red_dd4<A, B, false>(sx, sy, dr, di, lds + L_DD4);
      const PairScal<T> P = pair_scal<T, CPLX>(mx, my, kx, ky, sx, sy, dr, di);
      D = pair_decide<T, CPLX>(P, Sg.tol, Sg.gst, Sg.sec);
      o = pair_apply<T, CPLX, B, E, false>(0, xr, yr, xi, yi, nullptr, nullptr, nullptr, nullptr, P, D, 0, eSp, fSp, eSq,
                                           fSq, lds, lane);
      if (!o.err) {
        // (the update maximum and its check, as pair_apply makes them with
        // its own reduction)
        o.tG = red_max<T, B, false>(o.tG, lds + L_TG);
        if (((o.code == 2) || (o.code == -2) || (o.code == 3) || (o.code == -3)) && !isfinite_(o.tG))
          o.err = CPLX ? -630 : -580;
      }
    }
    if (!o.err && ((o.code != 1) || S.sG))
      store_rows<T, CPLX, B, E>(Gp, Gq, Gpi, Gqi, m, 0, lane, xr, yr, xi, yi);
    if (!D.err) {
      T tV = v_apply<T, CPLX, E>(D, vx, vy, vxi, vyi, nEV);
      if (vtrack && ((D.code == 2) || (D.code == -2)) && !((D.at == (T)0) && (!CPLX || (D.sat == (T)0)))) {
        tV = red_max<T, B>(tV, lds);
        o.tV = tV;
        if (!o.err && !isfinite_(tV))
          o.err = CPLX ? -638 : -588;
      }
    }
    if (S.sV || (!D.err && (D.code != 1) && (D.code != 3)))
      store_rows<T, CPLX, B, E>(Vp, Vq, Vpi, Vqi, n, 0, lane, vx, vy, vxi, vyi);
  }
  else if (!STREAM) {
    o = pair_step<T, CPLX, B, E>(m, n, xr, yr, xi, yi, Vp, Vq, Vpi, Vqi, eSp, fSp, eSq, fSq, S, lds, lane, D);
    if (!o.err && ((o.code != 1) || S.sG))
      store_rows<T, CPLX, B, E>(Gp, Gq, Gpi, Gqi, m, 0, lane, xr, yr, xi, yi);
  }
  else {
    typedef typename Acc<T>::A A;
    const int nc = (m + B * E - 1) / (B * E); // chunks of E rows per lane
    // The sums are taken in the same pass as the maxima, scaled by a
    // predicted exponent: the norm exponent each column was left with (every
    // step evaluates every column), moved by a pending rescaling.  Power-of-
    // two scaling commutes with every rounding of the sums as long as no
    // product or error term leaves the normal range under either exponent,
    // which holds when every nonzero element scaled by either exponent lies
    // in [2^-440, 2^100]; the sums are then rescaled to the max-exponent
    // scaling exactly, and otherwise taken again in a second pass.
    const T epx = eS[p], epy = eS[q];
    const bool pred = isfinite_(epx) && isfinite_(epy);
    const int kpx = pred ? (int)epx + S.sG : 0, kpy = pred ? (int)epy + S.sG : 0;
    T mx = (T)0, my = (T)0, nx = inf_<T>(), ny = inf_<T>(); // max and min nonzero |element|
    DD<A> sx{0, 0}, sy{0, 0}, dr{0, 0}, di{0, 0};
    constexpr int KN = Keep<T, CPLX, B, E>::N, KV = Keep<T, CPLX, B, E>::V;
    for (int c = 0; c < nc; ++c) {
      load_rows<T, CPLX, B, E>(Gp, Gq, Gpi, Gqi, m, c * E, S.sG, lane, xr, yr, xi, yi);
      if (c >= nc - KN)
        keep_rows<T, CPLX, B, E>(kbuf + (size_t)(nc - 1 - c) * KV * B, lane, xr, yr, xi, yi);
      T cx, cy;
      pair_maxima<T, CPLX, E>(xr, yr, xi, yi, E, cx, cy);
      mx = __builtin_fmax(mx, cx);
      my = __builtin_fmax(my, cy);
      if (pred) {
        pair_sums<T, CPLX, E>(xr, yr, xi, yi, E, kpx, kpy, sx, sy, dr, di);
#pragma unroll
        for (int j = 0; j < E; ++j) {
          nx = __builtin_fmin(nx, (xr[j] != (T)0) ? vabs(xr[j]) : inf_<T>());
          ny = __builtin_fmin(ny, (yr[j] != (T)0) ? vabs(yr[j]) : inf_<T>());
          if constexpr (CPLX) {
            nx = __builtin_fmin(nx, (xi[j] != (T)0) ? vabs(xi[j]) : inf_<T>());
            ny = __builtin_fmin(ny, (yi[j] != (T)0) ? vabs(yi[j]) : inf_<T>());
          }
        }
      }
    }
    red_max2<T, B>(mx, my, lds);
    const int kx = max_exp(mx), ky = max_exp(my);
    bool one = false;
    if (pred) {
      T mnx = -nx, mny = -ny;
      red_max2<T, B>(mnx, mny, lds);
      auto fits = [](const T mn, const T mxv, const int k, const int kp) {
        if (!(mxv > (T)0))
          return true; // a zero column: every sum is zero under any scaling
        const int en = frexp_exp(-mn) - 1;
        const int hi = (k > kp) ? k : kp;
        return (en - hi >= -440) && (k - kp <= 100);
      };
      one = fits(mnx, mx, kx, kpx) && fits(mny, my, ky, kpy);
    }
    if (one) {
      red_dd4<A, B>(sx, sy, dr, di, lds);
      const int dx = kpx - kx, dy = kpy - ky;
      sx = DD<A>{ldexp_(sx.s, 2 * dx), ldexp_(sx.c, 2 * dx)};
      sy = DD<A>{ldexp_(sy.s, 2 * dy), ldexp_(sy.c, 2 * dy)};
      dr = DD<A>{ldexp_(dr.s, dx + dy), ldexp_(dr.c, dx + dy)};
      di = DD<A>{ldexp_(di.s, dx + dy), ldexp_(di.c, dx + dy)};
    }
    else {
      sx = sy = dr = di = DD<A>{0, 0};
      for (int c = 0; c < nc; ++c) {
        load_rows<T, CPLX, B, E>(Gp, Gq, Gpi, Gqi, m, c * E, S.sG, lane, xr, yr, xi, yi);
        pair_sums<T, CPLX, E>(xr, yr, xi, yi, E, kx, ky, sx, sy, dr, di);
      }
      red_dd4<A, B>(sx, sy, dr, di, lds);
    }
    const PairScal<T> P = pair_scal<T, CPLX>(mx, my, kx, ky, sx, sy, dr, di);
    D = pair_decide<T, CPLX>(P, S.tol, S.gst, S.sec);
    // The update visits the chunks last to first: the ones the sums read
    // last are the likeliest to be still in the caches (the update of an
    // element does not depend on the others, so the order changes no bits;
    // the running maximum of the updated elements is order-independent for
    // the finite values a run produces).
    T tG = (T)0, tV = (T)0;
    // A pair left as it is (code 1, no pending rescaling) is not read again:
    // its update pass would change and store nothing (its maxima are 0).
    const bool idle = !D.err && (D.code == 1) && !S.sG;
    if (idle)
      o = pair_apply<T, CPLX, B, E, false>(0, xr, yr, xi, yi, Vp, Vq, Vpi, Vqi, P, D, S.vtrack, eSp, fSp, eSq, fSq,
                                           lds, lane, 0);
    for (int cc = 0; cc < (idle ? 0 : nc); ++cc) {
      const int c = nc - 1 - cc;
      if (cc < KN)
        kept_rows<T, CPLX, B, E>(kbuf + (size_t)cc * KV * B, lane, xr, yr, xi, yi);
      else
        load_rows<T, CPLX, B, E>(Gp, Gq, Gpi, Gqi, m, c * E, S.sG, lane, xr, yr, xi, yi);
      // the V pair is updated with the first chunk visited (whose call
      // reports the V maxima of a preset V)
      o = pair_apply<T, CPLX, B, E, false>(cc ? 0 : n, xr, yr, xi, yi, Vp, Vq, Vpi, Vqi, P, D, S.vtrack, eSp, fSp, eSq,
                                           fSq, lds, lane, E);
      if (o.err)
        break;
      if (cc == 0)
        tV = o.tV;
      tG = xmax(tG, o.tG);
      if ((o.code != 1) || S.sG)
        store_rows<T, CPLX, B, E>(Gp, Gq, Gpi, Gqi, m, c * E, lane, xr, yr, xi, yi);
    }
    if (!o.err) {
      tG = red_max<T, B>(tG, lds);
      o.tG = tG;
      o.tV = tV;
      if (((o.code == 2) || (o.code == -2) || (o.code == 3) || (o.code == -3)) && !isfinite_(tG))
        o.err = CPLX ? -630 : -580;
    }
  }
  if ((lane == 0) && vrec.code) {
    // the V update this pair owes, applied by k_vflush after the sweep (the
    // steps update V unless the decision failed)
    const long long ix = (long long)st_idx * gridDim.x + k;
    vrec.code[ix] = D.err ? 1 : D.code;
    vrec.c[ix] = D.c;
    vrec.at[ix] = D.at;
    if constexpr (CPLX)
      vrec.sat[ix] = D.sat;
    if ((st_idx == 0u) && (k == 0))
      st->vsweeps += 1u;
  }
  if (lane == 0) {
    // this pair's report, read by the next launch
    PlainSlot *const sl = slots + blockIdx.x;
    sl->err = o.err;
    if (!o.err) {
      eS[p] = eSp; fS[p] = fSp;
      eS[q] = eSq; fS[q] = fSq;
      sl->tG = (double)o.tG;
      sl->tV = (double)o.tV;
      sl->trans = (unsigned)o.trans;
      if (stats) {
        // |sin| of the rotation: c |tan|, the complex tangent being at + i sat
        T ta = vabs(D.at);
        if constexpr (CPLX)
          ta = sqrt_(fma_(D.sat, D.sat, ta * ta));
        const T sn = D.trans ? D.c * ta : (T)0;
        sl->cosv = (double)D.cosv;
        sl->sinv = (double)sn;
      }
    }
  }
}

// ----------------------------------------------------------------- epilogue
template <typename T>
__global__ __launch_bounds__(256) void k_epilogue(const int m, const int n, T *__restrict__ Gr, T *__restrict__ Gi,
                                                  const long long ldG, T *__restrict__ Vr, T *__restrict__ Vi,
                                                  const long long ldV, T *__restrict__ eS, T *__restrict__ fS,
                                                  const PlainState<T> *__restrict__ st, const unsigned swp)
{
  // dvjsvd.c:604-631: G_j <- G_j / fS_j * 2^-eS_j, eS_j -= sT; V *= 2^-sV
  if (st->status || st->zero || !(st->sweeps < swp))
    return;
  const int j = blockIdx.x;
  const T f_ = fS[j], s_ = -eS[j];
  for (int i = threadIdx.x; i < m; i += blockDim.x) {
    Gr[j * ldG + i] = scalef(Gr[j * ldG + i] / f_, s_);
    if (Gi)
      Gi[j * ldG + i] = scalef(Gi[j * ldG + i] / f_, s_);
  }
  if (st->sV) {
    const T sv = (T)(-st->sV);
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
      Vr[j * ldV + i] = scalef(Vr[j * ldV + i], sv);
      if (Vi)
        Vi[j * ldV + i] = scalef(Vi[j * ldV + i], sv);
    }
  }
  __syncthreads();
  if (threadIdx.x == 0)
    eS[j] -= (T)st->sT;
}

// ------------------------------------------------------------------ driver
template <typename T, bool CPLX, int B, int E, bool STREAM = false>
static int launch_steps(const PlainArgs<T> &a, PlainState<T> *st, unsigned r, const PlainSlot *prev, PlainSlot *slots,
                        const unsigned *row, unsigned s, const T tol, const T gst, const VRec<T> &vrec,
                        hipStream_t strm)
{
  // with deferred V updates the steps see V columns of length 0
  const int nv = vrec.code ? 0 : (int)(a.nv ? a.nv : a.n);
  hipLaunchKernelGGL((k_step<T, CPLX, B, E, STREAM>), dim3(a.n / 2), dim3(B), 0, strm, (int)a.m, nv, a.Gr, a.Gi,
                     a.ldG, a.Vr, a.Vi, a.ldV, a.eS, a.fS, row, s, a.stp, a.swp, tol, gst, a.vpreset, a.stats, a.sec,
                     st, r, prev, slots, vrec);
  return 0;
}

template <typename T, bool CPLX>
using StepFn = int (*)(const PlainArgs<T> &, PlainState<T> *, unsigned, const PlainSlot *, PlainSlot *,
                       const unsigned *, unsigned, T, T, const VRec<T> &, hipStream_t);

template <typename T, bool CPLX>
static StepFn<T, CPLX> pick_step(long long m, int &B_out)
{
  // smallest group that holds the column pair in registers with <= 8
  // elements per lane, then 512 lanes (two waves per SIMD, 256 registers per
  // lane) with 8 or 16; longer columns are streamed (a 1024-lane group would
  // cap the registers at 128 and spill the pair): real ones through 512
  // lanes in chunks of 4 rows per lane, complex ones through 256 lanes in
  // chunks of 2 (116 registers, four workgroups per CU; on an 8192 x 2048
  // complex matrix 21% faster than 512 lanes with chunks of 4, which need
  // 154 registers and fit one workgroup per CU).  Complex columns of 2049 to
  // 4096 rows stream too: in registers they need 228 per lane (one
  // 512-lane workgroup per CU), 6% slower on 4096 x 1024; up to 2048 rows
  // the registers win (by 23% on 2048 x 512).
  if (m <= 64) { B_out = 64; return &launch_steps<T, CPLX, 64, 1>; }
  if (m <= 128) { B_out = 64; return &launch_steps<T, CPLX, 64, 2>; }
  if (m <= 256) { B_out = 64; return &launch_steps<T, CPLX, 64, 4>; }
  if (m <= 512) { B_out = 64; return &launch_steps<T, CPLX, 64, 8>; }
  if (m <= 1024) { B_out = 256; return &launch_steps<T, CPLX, 256, 4>; }
  if (m <= 2048) { B_out = 256; return &launch_steps<T, CPLX, 256, 8>; }
  if constexpr (!CPLX) {
    if (m <= 4096) { B_out = 512; return &launch_steps<T, CPLX, 512, 8>; }
    if (m <= 8192) { B_out = 512; return &launch_steps<T, CPLX, 512, 16>; }
  }
  if constexpr (CPLX) {
    if (m <= 0x7fffffffLL / 2) { B_out = 256; return &launch_steps<T, CPLX, 256, 2, true>; }
  }
  if (m <= 0x7fffffffLL / 2) { B_out = 512; return &launch_steps<T, CPLX, 512, 4, true>; }
  return nullptr;
}

template <typename T> static T ldexp_host(T x, int e) { return std::ldexp(x, e); }

// Shape of the deferred V updates: R rows of V per workgroup (the most, up
// to 16, that keep R n elements within the LDS budget and at least two
// workgroups per CU), nt lanes, PM pairs per lane.  False when a row of V
// does not fit the budget or a step has more than 4096 pairs: the steps then
// update V themselves.
struct VFlushPlan {
  int R, nt, pm;
};

template <typename T, bool CPLX> static bool vflush_plan(long long n, long long nv, int cus, VFlushPlan &f)
{
  const long long row = n * (long long)sizeof(T) * (CPLX ? 2 : 1);
  const long long np = n / 2;
  if ((row > VF_LDS_MAX) || (np > 4096) || (np < 1))
    return false;
  f.nt = (int)std::min<long long>(1024, (np + 63) / 64 * 64);
  f.pm = (int)((np + f.nt - 1) / f.nt);
  if (f.pm == 3)
    f.pm = 4;
  f.R = 1;
  while ((f.R < 16) && (2 * f.R * row <= VF_LDS_MAX) && (nv / (2 * f.R) >= 2LL * cus))
    f.R *= 2;
  return true;
}

template <typename T, bool CPLX>
static int launch_vflush(const PlainArgs<T> &a, const VFlushPlan &f, const VRec<T> &vrec, const PlainState<T> *st,
                         unsigned sw, hipStream_t s)
{
  const long long nv = a.nv ? a.nv : a.n;
  const unsigned grid = (unsigned)((nv + f.R - 1) / f.R);
  const size_t lds = (size_t)f.R * (size_t)a.n * sizeof(T) * (CPLX ? 2 : 1);
#define VJ_VFLUSH(PM)                                                                                              \
  hipLaunchKernelGGL((k_vflush<T, CPLX, PM>), dim3(grid), dim3(f.nt), lds, s, (int)nv, (int)a.n, f.R, a.Vr, a.Vi, \
                     a.ldV, a.js, a.stp, vrec, st, sw, a.sec)
  if (f.pm == 1)
    VJ_VFLUSH(1);
  else if (f.pm == 2)
    VJ_VFLUSH(2);
  else
    VJ_VFLUSH(4);
#undef VJ_VFLUSH
  return check_launch();
}

static __global__ void k_zero_words(unsigned *p, size_t words)
{
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < words; i += (size_t)gridDim.x * blockDim.x)
    p[i] = 0u;
}

int zero_words(void *p, size_t bytes, hipStream_t s)
{
  const size_t words = (bytes + 3u) / 4u;
  const unsigned grid = (unsigned)std::min<size_t>((words + 255u) / 256u, 1024u);
  hipLaunchKernelGGL(k_zero_words, dim3(grid ? grid : 1u), dim3(256), 0, s, (unsigned *)p, words);
  return check_launch();
}

template <typename T, bool CPLX>
int plain_svd(const PlainArgs<T> &a, PlainRun *run, hipStream_t s, const PlainWork<T> *work)
{
  if (a.n == 0)
    return 0;
  int B = 0;
  StepFn<T, CPLX> step = pick_step<T, CPLX>(a.m, B);
  if (!step)
    return set_error(-1, "m = %lld exceeds the largest supported column length", a.m);
  // control block and per-pair slots (two steps' worth: a step reads the
  // previous step's slots while it writes its own), released in stream order
  // on return, or the caller's workspace; a step reads the previous step's
  // slots only after the first step of a sweep, so the slots start unset
  Staging sg(s);
  PlainState<T> *st = nullptr;
  PlainSlot *slots = nullptr;
  const size_t np = (size_t)(a.n / 2);
  if (work) {
    st = work->st;
    slots = work->slots;
  }
  else {
    if (int rc = sg.alloc(&st, 1))
      return rc;
    if (int rc = sg.alloc(&slots, 2 * np))
      return rc;
  }
  if (int rc = zero_words(st, sizeof(PlainState<T>), s))
    return rc;
  const int mx_blocks = 1024;
  hipLaunchKernelGGL(k_maxabs<T>, dim3(mx_blocks), dim3(256), 0, s, (int)a.m, (int)a.n, a.Gr, a.ldG, &st->MG_u);
  if (CPLX)
    hipLaunchKernelGGL(k_maxabs<T>, dim3(mx_blocks), dim3(256), 0, s, (int)a.m, (int)a.n, a.Gi, a.ldG, &st->MG_u);
  const long long nv = a.nv ? a.nv : a.n;
  if (a.vpreset) {
    hipLaunchKernelGGL(k_maxabs<T>, dim3(mx_blocks), dim3(256), 0, s, (int)nv, (int)a.n, a.Vr, a.ldV, &st->MV_u);
    if (CPLX)
      hipLaunchKernelGGL(k_maxabs<T>, dim3(mx_blocks), dim3(256), 0, s, (int)nv, (int)a.n, a.Vi, a.ldV, &st->MV_u);
  }
  hipLaunchKernelGGL((k_setup<T, CPLX>), dim3(1), dim3(64), 0, s, (int)a.m, st, a.vpreset);
  hipLaunchKernelGGL(k_init_cols<T>, dim3((unsigned)a.n), dim3(256), 0, s, (int)a.m, (int)nv, a.Gr, a.Gi, a.ldG,
                     a.Vr, a.Vi, a.ldV, a.eS, a.fS, st, a.vpreset);
  if (int rc = check_launch())
    return rc;
  const T tol = (a.tol > (T)0) ? a.tol : host_tol<T>(a.m);
  const T gst = host_gst<T>(tol);
  // V := I and no caller's workspace: the V updates of each sweep are
  // deferred to one pass over V after it (same bits)
  VRec<T> vrec{nullptr, nullptr, nullptr, nullptr};
  VFlushPlan vfp{0, 0, 0};
  if (!a.vpreset && !work) {
    int dev = 0, cus = 0;
    VJ_HIP(hipGetDevice(&dev));
    VJ_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    if (vflush_plan<T, CPLX>(a.n, nv, cus, vfp)) {
      const size_t nrec = (size_t)a.stp * np;
      if (int rc = sg.alloc(&vrec.code, nrec))
        return rc;
      if (int rc = sg.alloc(&vrec.c, nrec))
        return rc;
      if (int rc = sg.alloc(&vrec.at, nrec))
        return rc;
      if (CPLX) {
        if (int rc = sg.alloc(&vrec.sat, nrec))
          return rc;
      }
      const size_t lds = (size_t)vfp.R * (size_t)a.n * sizeof(T) * (CPLX ? 2 : 1);
      if (lds > 65536u) {
        const void *fn = (vfp.pm == 1) ? (const void *)k_vflush<T, CPLX, 1>
                         : (vfp.pm == 2) ? (const void *)k_vflush<T, CPLX, 2> : (const void *)k_vflush<T, CPLX, 4>;
        VJ_HIP(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
      }
    }
  }
  // Sweeps are enqueued ahead of the convergence check: the host reads the
  // device flag of sweep k-1 after the first steps of sweep k have been
  // queued (enough work to cover the wait and the queueing of the rest), and
  // queues the rest of sweep k only when the run goes on.  With a caller's
  // workspace the flag of each sweep is read synchronously instead.
  PlainFlags *hf = nullptr;
  hipEvent_t ev[2] = {nullptr, nullptr};
  if (!work) {
    VJ_HIP(hipHostMalloc((void **)&hf, 2 * sizeof(PlainFlags), hipHostMallocDefault));
    VJ_HIP(hipEventCreateWithFlags(&ev[0], hipEventDisableTiming));
    VJ_HIP(hipEventCreateWithFlags(&ev[1], hipEventDisableTiming));
  }
  int rc = 0;
  unsigned launched = 0;
  hipEvent_t tev[2] = {nullptr, nullptr};
  const bool timed = run && !work;
  if (timed && (hipEventCreate(&tev[0]) != hipSuccess || hipEventCreate(&tev[1]) != hipSuccess ||
                hipEventRecord(tev[0], s) != hipSuccess))
    rc = set_error(VJ_EHIP, "event creation failed");
  // ctl[r] is the control state before the next launch; each launch that
  // folds a step moves it to ctl[r ^ 1]
  unsigned r = 0;
  const unsigned head = std::min(a.stp, 128u); // steps queued before the previous sweep's flag is read
  for (unsigned sw = 0; !rc && sw < a.swp; ++sw) {
    bool ended = false;
    for (unsigned k = 0; k < a.stp; ++k) {
      if (!work && (sw >= 1) && (k == head)) {
        ++launched; // (partly queued)
        if (hipEventSynchronize(ev[(sw - 1) & 1u]) != hipSuccess) {
          rc = set_error(VJ_EHIP, "event sync failed");
          break;
        }
        if (hf[(sw - 1) & 1u].done) {
          // the queued steps return at once; their launches that fold keep
          // the control state's parity (done launches copy ctl[r] forward)
          ended = true;
          break;
        }
        --launched;
      }
      step(a, st, r, slots + ((k + 1u) & 1u) * np, slots + (k & 1u) * np, a.js + (size_t)k * a.n, k, tol, gst, vrec,
           s);
      if (k)
        r ^= 1u;
    }
    if (rc || ended)
      break;
    hipLaunchKernelGGL((k_fold<T, CPLX>), dim3(1), dim3(256), 0, s, st, slots + ((a.stp - 1u) & 1u) * np, (int)np, r,
                       a.stp, a.swp, a.stats);
    r ^= 1u;
    if ((rc = check_launch()))
      break;
    if (vrec.code && (rc = launch_vflush<T, CPLX>(a, vfp, vrec, st, sw, s)))
      break;
    ++launched;
    if (work) {
      PlainFlags f{};
      if (hipMemcpyAsync(&f, &st->done, sizeof(PlainFlags), hipMemcpyDeviceToHost, s) != hipSuccess ||
          hipStreamSynchronize(s) != hipSuccess) {
        rc = set_error(VJ_EHIP, "flag copy failed");
        break;
      }
      if (f.done)
        break;
      continue;
    }
    if (hipMemcpyAsync(&hf[sw & 1u], &st->done, sizeof(PlainFlags), hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipEventRecord(ev[sw & 1u], s) != hipSuccess) {
      rc = set_error(VJ_EHIP, "flag copy failed");
      break;
    }
    if ((sw >= 1) && (a.stp <= head)) {
      // a sweep of at most `head` steps: its flag read after the whole sweep
      if (hipEventSynchronize(ev[(sw - 1) & 1u]) != hipSuccess) {
        rc = set_error(VJ_EHIP, "event sync failed");
        break;
      }
      if (hf[(sw - 1) & 1u].done)
        break;
    }
  }
  if (!rc && !a.raw) {
    hipLaunchKernelGGL(k_epilogue<T>, dim3((unsigned)a.n), dim3(256), 0, s, (int)a.m, (int)nv, a.Gr, a.Gi, a.ldG,
                       a.Vr, a.Vi, a.ldV, a.eS, a.fS, st, a.swp);
    rc = check_launch();
  }
  if (!rc && timed && hipEventRecord(tev[1], s) != hipSuccess)
    rc = set_error(VJ_EHIP, "event record failed");
  PlainState<T> h;
  if (!rc) {
    if (hipMemcpyAsync(&h, st, sizeof(h), hipMemcpyDeviceToHost, s) != hipSuccess || hipStreamSynchronize(s) != hipSuccess)
      rc = set_error(VJ_EHIP, "state copy failed");
  }
  float steps_ms = 0.0f;
  if (!rc && timed && hipEventElapsedTime(&steps_ms, tev[0], tev[1]) != hipSuccess)
    steps_ms = -1.0f;
  for (int i = 0; i < 2; ++i) {
    if (tev[i])
      (void)hipEventDestroy(tev[i]);
    if (ev[i])
      (void)hipEventDestroy(ev[i]);
  }
  if (hf)
    (void)hipHostFree(hf);
  if (rc)
    return rc;
  if (run) {
    run->sweeps = h.sweeps;
    run->sT = h.sT;
    run->sV = h.sV;
    run->launched_sweeps = launched;
    for (int i = 0; i < 64; ++i)
      run->hist[i] = h.hist[i];
    run->MG0 = (double)h.MG0;
    run->MG_scaled = (double)ldexp_host(h.MG0, h.sG_init);
    run->MG_final = (double)h.ctl[r].MG;
    run->eM0 = h.eM0;
    run->sR0 = h.sR0;
    run->sN0 = h.sG_init;
    run->zero = h.zero;
    run->status = h.status;
    run->nev = h.nev;
    for (unsigned i = 0; i < h.nev && i < 32u; ++i) {
      run->ev[i].sw = h.ev[i].sw;
      run->ev[i].st = h.ev[i].st;
      run->ev[i].eM = h.ev[i].eM;
      run->ev[i].sR = h.ev[i].sR;
      run->ev[i].sN = h.ev[i].sN;
      run->ev[i].M = (double)h.ev[i].M;
    }
    run->err_sw = h.err_sw;
    run->err_st = h.err_st;
    run->steps_ms = steps_ms;
    run->last_cos = h.last_cos;
    run->last_sin = h.last_sin;
    run->sG_pending = h.ctl[r].sG;
    run->sV_pending = h.ctl[r].sV;
  }
  if (h.status)
    return h.status;
  if (h.zero)
    return 0;
  return (int)h.sweeps;
}

template int plain_svd<double, false>(const PlainArgs<double> &, PlainRun *, hipStream_t, const PlainWork<double> *);
template int plain_svd<double, true>(const PlainArgs<double> &, PlainRun *, hipStream_t, const PlainWork<double> *);
template int plain_svd<float, false>(const PlainArgs<float> &, PlainRun *, hipStream_t, const PlainWork<float> *);
template int plain_svd<float, true>(const PlainArgs<float> &, PlainRun *, hipStream_t, const PlainWork<float> *);

} // namespace vj
