// Modulus ("MM", pjs id 4) pairing of the strategy tables: the one
// definition shared by the host table builder (pjs) and the batched kernel,
// which computes its pairs instead of reading a table.
#pragma once

#include <hip/hip_runtime.h>

namespace vj {

// Pair k (0 <= k < n/2) of step s (0 <= s < n) for even n >= 2: the pairs
// {i, j} with i + j = s (mod n), i < j, in increasing i, then for even s the
// two fixed points of i -> s - i, {s/2, s/2 + n/2}.  Over the n steps every
// pair meets once, the n/2 "diameter" pairs twice.  The smaller index comes
// first.
__host__ __device__ inline void mod_pair(const unsigned n, const unsigned s, const unsigned k, unsigned &p, unsigned &q)
{
  const unsigned g1 = (s + 1u) >> 1u;                                              // {i, s - i}
  const unsigned g2 = ((s + n) & 1u) ? ((n - s - 1u) >> 1u) : (((n - s) >> 1u) - 1u); // {i, s + n - i}
  unsigned a, b;
  if (k < g1) {
    a = k;
    b = s - k;
  }
  else if (k < g1 + g2) {
    a = s + 1u + (k - g1);
    b = s + n - a;
  }
  else {
    a = s >> 1u;
    b = a + (n >> 1u);
  }
  p = (a < b) ? a : b;
  q = (a < b) ? b : a;
}

} // namespace vj
