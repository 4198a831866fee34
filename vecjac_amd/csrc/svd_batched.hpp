// Arguments of the batched SVD kernel (all pointers are device memory).
#pragma once

#include <hip/hip_runtime.h>

namespace vj {

template <typename T> struct BatchArgs {
  int nb;                         // number of matrices
  const long long *m, *n;         // shapes (n even, n <= m)
  const long long *offG, *offV;   // element offsets of G_b (ld m_b) and V_b (ld n_b)
  const long long *offS;          // offsets of eS_b / fS_b
  const int *order;               // processing order (decreasing cost)
  T *Gr, *Gi, *Vr, *Vi, *eS, *fS; // data (Gi/Vi null for real data)
  long long *info;                // per matrix: sweeps, or a negative code
  unsigned long long *transforms; // per matrix: transformations (may be null)
  unsigned long long *evals;      // per matrix: pair evaluations (may be null)
  unsigned long long *xfer;       // per matrix: elements of G and V read or written in memory (may be null)
  unsigned *queue;                // work counter, zero on entry
  unsigned swp;                   // sweep limit
  int lds_bytes;                  // LDS-resident form: bytes of G per workgroup (0: global)
  int group_lg;                   // lane-group form: 2^group_lg lanes per pair (4, 3 or 2)
  int ord;                        // ordering of the pairs: pjs id 4 (modulus) or 6 (block modulus)
  int vpreset;                    // V holds the starting matrices (else V := I)
};

// Shapes of the lane-group form that are built: 16 lanes per pair with any
// waves per workgroup; 8 lanes with up to 4 waves (complex up to 64 rows);
// 4 lanes with up to 2 waves, real, up to 64 rows.  rows: the largest m.
constexpr bool group_shape_ok(bool cplx, int rows, int lg, int waves)
{
  return (lg == 4) || (lg == 3 && waves <= 4 && (!cplx || rows <= 64)) ||
         (lg == 2 && waves <= 2 && !cplx && rows <= 64);
}
constexpr int group_max_rows(bool cplx, int lg)
{
  return cplx ? 64 : ((lg == 2) ? 64 : 128);
}

// One launch over a.nb matrices of a.order: grp selects the short-column
// form (lane groups of 16, m <= 128 real / 64 complex), else the wave form;
// a.lds_bytes > 0 (short-column form only) keeps each matrix's G in LDS for
// all its sweeps.
template <typename T, bool CPLX>
int batched_svd(const BatchArgs<T> &a, int waves, int max_m, int grid_cap, bool grp, hipStream_t s);

// LDS bytes a matrix needs in the LDS-resident form: G (m x n), both planes
// for complex data (V stays in device memory).
inline long long lds_matrix_bytes(long long m, long long n, bool cplx, int tsize)
{
  return m * n * (long long)tsize * (cplx ? 2 : 1);
}

// Largest dynamic LDS a workgroup of the LDS-resident form may take (the
// 160 KiB of a gfx950 CU less the kernel's own static LDS).
constexpr int LDS_FORM_MAX_BYTES = 152 * 1024;

// Rows up to which a matrix of a batch takes the short-column form.
inline int short_rows(bool cplx) { return cplx ? 64 : 128; }

} // namespace vj
