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
  unsigned *queue;                // work counter, zero on entry
  unsigned swp;                   // sweep limit
};

template <typename T, bool CPLX> int batched_svd(const BatchArgs<T> &a, int waves, int max_m, int grid_cap, hipStream_t s);

} // namespace vj
