// Shared declarations of the SVD drivers (plain and batched).
#pragma once

#include "capi_common.hpp"
#include "svd_pair.hpp"

#include <cmath>

namespace vj {

// Arguments of one plain SVD on device memory (column-major, split re/im
// planes for the complex drivers; Gi/Vi are null for real data).
template <typename T> struct PlainArgs {
  long long m, n;
  T *Gr, *Gi;
  long long ldG;
  T *Vr, *Vi;
  long long ldV;
  T *eS, *fS;
  const unsigned *js; // device step table, stp rows of n entries
  unsigned stp, swp;
  int vpreset;        // iwork[0] != 0: V holds a starting matrix
  // options of the LAPACK-style drivers (zero: the reference's behaviour)
  long long nv;       // rows of V (0: n); a preset nv x n V has the rotations applied
  T tol;              // convergence threshold (0: sqrt(m) eps/2)
  int raw;            // no epilogue: G, V and eS stay in their scaled working form
  int stats;          // record max |cos| and max |sin| of the pairs of every sweep
  int sec;            // the reference's USE_SECANTS form: secants, rotations divide by them
};

// The first two words are copied to the host once per sweep.
struct PlainFlags {
  int done;
  int status;
};

// The part of the control state that every workgroup of a step reads.  It
// is double-buffered: the workgroups of a step read ctl[r] and fold the
// previous step's slots into their own copy, and workgroup 0 stores that
// copy in ctl[r ^ 1] for the next launch.
template <typename T> struct PlainCtl {
  T MG, MV;                      // running max-|.| bounds (dvjsvd.c:78, 88)
  int sG, sV;                    // rescaling to apply during the next step
};

// Device-resident control block of one plain SVD.
template <typename T> struct PlainState {
  int done;                      // converged, failed or out of sweeps
  int status;                    // 0 or the reference's negative code
  int zero;                      // G == 0: early return with 0 sweeps
  int ef_init;                   // inputs accepted: V/eS/fS initialised
  PlainCtl<T> ctl[2];
  unsigned long long MG_u, MV_u; // initial max-|.| (atomicMax on bit patterns)
  int NRM_EXP;                   // exponent of max/(2m) (dvjsvd.c:120-123)
  int sT, sV;                    // accumulated power-of-two shifts
  int sG_init, sV_init;          // initial scalings
  unsigned sweeps;               // completed sweeps with transformations
  unsigned vsweeps;              // sweeps whose steps ran (deferred V updates)
  unsigned long long swt;        // transformations in the current sweep
  unsigned long long hist[64];   // transformations per sweep (trace)
  // trace of the scalings (the reference's JTRACE lines, dvjsvd.c:70-635)
  T MG0;                         // max |G| on entry
  int eM0, sR0;                  // its exponent and the rotation headroom
  unsigned nev;                  // rescalings logged in ev[]
  struct {
    unsigned sw, st;             // sweep and step the rescaling precedes
    int eM, sR, sN;              // exponent of MG, headroom, applied shift
    T M;                         // MG after the shift
  } ev[32];
  unsigned err_sw, err_st;       // sweep and step of a failure
  // pair statistics (stats != 0): maxima of |cos(x, y)| over the pairs
  // examined and of |sin| over the rotations applied, in the current sweep
  // and in the last one completed
  double cos_max, sin_max, last_cos, last_sin;
};

// What one workgroup of a step reports: one slot per pair.  The next launch
// (or the end-of-sweep fold) reduces all slots of the step; the kernel
// boundary orders the stores before those loads.
struct PlainSlot {
  double tG, tV;                 // max |updated G / V element|
  double cosv, sinv;             // pair statistics (stats != 0)
  unsigned trans;                // 1 if transformed
  int err;                       // 0, or the reference's negative code
};

struct PlainRun {
  unsigned sweeps, launched_sweeps;
  int sT, sV;
  unsigned long long hist[64];
  // trace (the reference's JTRACE output)
  double MG0, MG_scaled, MG_final;
  int eM0, sR0, sN0;
  int zero, status;
  unsigned nev;
  struct {
    unsigned sw, st;
    int eM, sR, sN;
    double M;
  } ev[32];
  unsigned err_sw, err_st;
  float steps_ms;                // GPU time of the sweeps (HIP events)
  double last_cos, last_sin;     // pair statistics of the last sweep run
  int sG_pending, sV_pending;    // shifts counted in sT / sV but not yet applied
};

// Device workspace handed in by the caller (the allocation-free entry
// point): the control block and 2 * (n/2) slots.  With it the driver makes no
// allocation of its own: the control block is cleared by a fill kernel, and
// the convergence flag is read back synchronously after each sweep (no
// pinned host flags, no events), so the sweeps are not queued one ahead.
template <typename T> struct PlainWork {
  PlainState<T> *st;
  PlainSlot *slots;
};

template <typename T, bool CPLX>
int plain_svd(const PlainArgs<T> &a, PlainRun *run, hipStream_t s, const PlainWork<T> *work = nullptr);

// Clears `bytes` (rounded up to whole 32-bit words) of device memory with a
// fill kernel on stream s, in place of hipMemsetAsync.
int zero_words(void *p, size_t bytes, hipStream_t s);

// Host-side thresholds, evaluated exactly as the reference does
// (dvjsvd.c:183-184): tol = sqrt(m) * eps/2, gst = tol * 2^FIN_EXP.
template <typename T> inline T host_tol(long long m) { return std::sqrt((T)m) * std::ldexp(Lim<T>::EPS, -1); }
template <typename T> inline T host_gst(T tol) { return std::ldexp(tol, (int)Lim<T>::FIN_EXP); }

} // namespace vj
