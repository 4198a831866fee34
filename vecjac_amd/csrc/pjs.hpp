// Pairings of the strategy tables that the batched kernel computes instead
// of reading a table (modulus, pjs id 4; block modulus, pjs id 6): the one
// definition shared by the host table builder (pjs) and the kernel.
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

// Block modulus ordering (pjs id 6, an extension of this library): the n
// columns form nb = n/2 blocks {2I, 2I+1}, and block round s (0 <= s < nb)
// meets block I with block J = s - I (mod nb).  A round is two steps: for
// I < J, step 2s pairs (2I, 2J), (2I+1, 2J+1) and step 2s+1 pairs
// (2I, 2J+1), (2I+1, 2J); a block met by itself (2I = s mod nb) pairs its
// own two columns at both steps.  Every cross-block pair meets once per
// sweep of n steps, every in-block pair twice; each step is a permutation
// of 0..n-1.  Because the four columns of a block pair take part in no other
// pair during the round, the batched kernel can keep them in registers for
// both steps (one load and one store per column per round).
//
// Units of round s: the block pairs {I, s-I} (I < s-I), then those with
// I + J = s + nb, then the blocks met by themselves (s/2 for even s, then
// (s+nb)/2 for even s+nb).
__host__ __device__ inline unsigned br_units(const unsigned nb, const unsigned s)
{
  return ((s + 1u) >> 1u) + ((nb - 1u - s) >> 1u) + ((s & 1u) ^ 1u) + (((s + nb) & 1u) ^ 1u);
}

// Unit u of round s: blocks I <= J (I == J for a block met by itself).
__host__ __device__ inline void br_unit(const unsigned nb, const unsigned s, const unsigned u, unsigned &I, unsigned &J)
{
  const unsigned g1 = (s + 1u) >> 1u, g2 = (nb - 1u - s) >> 1u;
  if (u < g1) {
    I = u;
    J = s - u;
  }
  else if (u < g1 + g2) {
    I = s + 1u + (u - g1);
    J = s + nb - I;
  }
  else {
    const unsigned d = u - g1 - g2;
    I = J = ((d == 0u) && !(s & 1u)) ? (s >> 1u) : ((s + nb) >> 1u);
  }
}

// Pair k (0 <= k < n/2) of step st (0 <= st < n) for even n >= 2: the two
// pairs of each block pair of the round in unit order, then the in-block
// pairs.  The smaller index comes first.
__host__ __device__ inline void br_pair(const unsigned n, const unsigned st, const unsigned k, unsigned &p, unsigned &q)
{
  const unsigned nb = n >> 1u, s = st >> 1u, h = st & 1u;
  const unsigned g = ((s + 1u) >> 1u) + ((nb - 1u - s) >> 1u);
  unsigned I, J;
  if (k < 2u * g) {
    br_unit(nb, s, k >> 1u, I, J);
    const unsigned x = k & 1u;
    p = 2u * I + x;
    q = 2u * J + (h ? (x ^ 1u) : x);
  }
  else {
    br_unit(nb, s, g + (k - 2u * g), I, J);
    p = 2u * I;
    q = p + 1u;
  }
}

// Pair k of step st of ordering ord (4: modulus, 6: block modulus).
__host__ __device__ inline void step_pair(const int ord, const unsigned n, const unsigned st, const unsigned k,
                                          unsigned &p, unsigned &q)
{
  if (ord == 6)
    br_pair(n, st, k, p, q);
  else
    mod_pair(n, st, k, p, q);
}

} // namespace vj
