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
  unsigned *queue;                // work counter, zero on entry
  unsigned swp;                   // sweep limit
};

// One launch over a.nb matrices of a.order: grp selects the short-column
// form (lane groups of 16, m <= 128 real / 64 complex), else the wave form.
template <typename T, bool CPLX>
int batched_svd(const BatchArgs<T> &a, int waves, int max_m, int grid_cap, bool grp, hipStream_t s);

// Rows up to which a matrix of a batch takes the short-column form.
inline int short_rows(bool cplx) { return cplx ? 64 : 128; }

} // namespace vj
