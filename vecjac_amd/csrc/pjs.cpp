// Parallel Jacobi strategy tables (host side), the input of every SVD driver.
//
// Replaces pjs() (pjs.c:7-65, pjs.h:9-18), whose orderings come from the
// external libpvn generator.  Row st of the stp x n table lists n/2
// disjoint pairs (r[2k], r[2k+1]); the SVD kernels read the table straight
// from device memory, one row per step.  Two strategies, keyed by the
// reference's ids:
//   id 2 (PJS_ME, n-1 steps): round-robin ("circle") tournament: each
//        unordered pair meets exactly once per sweep;
//   id 4 (PJS_MM, n steps): modulus ordering: pair {i,j} meets in step
//        (i+j) mod n, and for even steps the two indices with 2i == step
//        (mod n) pair with each other, so each row is a perfect matching and
//        the n/2 "diameter" pairs {i, i+n/2} meet twice per sweep.
// Every row is a permutation of 0..n-1 (needed for race-free column
// updates); each pair is written with its smaller index first.
#include "pjs.hpp"

#include <stdint.h>
#include <stdlib.h>

namespace {

inline void put(unsigned *r, unsigned k, unsigned a, unsigned b)
{
  r[2u * k] = (a < b) ? a : b;
  r[2u * k + 1u] = (a < b) ? b : a;
}

void round_robin(unsigned n, unsigned *tab)
{
  for (unsigned s = 0u; s + 1u < n; ++s) {
    unsigned *const r = tab + (size_t)s * n;
    for (unsigned k = 0u; k < (n >> 1u); ++k)
      vj::rr_pair(n, s, k, r[2u * k], r[2u * k + 1u]);
  }
}

void modulus(unsigned n, unsigned *tab)
{
  for (unsigned s = 0u; s < n; ++s) {
    unsigned *const r = tab + (size_t)s * n;
    unsigned k = 0u;
    for (unsigned i = 0u; i < n; ++i) {
      const unsigned j = (s + n - i) % n;
      if (i < j)
        put(r, k++, i, j);
    }
    if (!(s & 1u)) // the two fixed points of i -> s - i
      put(r, k++, s >> 1u, (s >> 1u) + (n >> 1u));
  }
}

} // namespace

extern "C" unsigned *pjs(const long id, const unsigned n, unsigned *stp)
{
  *stp = 0u;
  if ((n <= 1u) || (n & 1u))
    return (unsigned *)NULL;
  unsigned s = 0u;
  switch (id) {
  case 2l:
    s = n - 1u;
    break;
  case 4l:
    s = n;
    break;
  default:
    return (unsigned *)NULL;
  }
  unsigned *const tab = (unsigned *)malloc((size_t)n * s * sizeof(unsigned));
  if (!tab)
    return (unsigned *)NULL;
  if (id == 2l)
    round_robin(n, tab);
  else
    modulus(n, tab);
  *stp = s;
  return tab;
}

extern "C" void pjs_free(unsigned *tab) { free(tab); }
