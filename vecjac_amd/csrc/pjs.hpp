// Round-robin ("circle method") pairing of the strategy tables with id 2:
// the one definition shared by the host table builder (pjs) and the batched
// kernel, which computes its pairs instead of reading a table.
#pragma once

#include <hip/hip_runtime.h>

namespace vj {

// Pair k (0 <= k < n/2) of step s (0 <= s < n-1) for even n >= 2: column n-1
// stays fixed and meets column s; the others pair up symmetrically around s
// on the circle of the n-1 remaining columns.  The smaller index comes first.
__host__ __device__ inline void rr_pair(const unsigned n, const unsigned s, const unsigned k, unsigned &p, unsigned &q)
{
  const unsigned m = n - 1u;
  unsigned a = s, b = m;
  if (k) {
    a = s + k;
    a = (a >= m) ? (a - m) : a;
    b = s + m - k;
    b = (b >= m) ? (b - m) : b;
  }
  p = (a < b) ? a : b;
  q = (a < b) ? b : a;
}

} // namespace vj
