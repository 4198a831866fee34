// Parallel Jacobi strategy tables (host side), the input of every SVD driver.
//
// Replaces pjs() (pjs.c:7-65, pjs.h:9-18), whose orderings come from the
// external libpvn generator.  Row st of the stp x n table lists n/2
// disjoint pairs (r[2k], r[2k+1]); the SVD kernels read the table straight
// from device memory, one row per step.  Two strategies, keyed by the
// reference's ids:
//   id 2 (PJS_ME, n-1 steps): the block-recursive ordering of Mantharam and
//        Eberlein: the index set is halved, its two halves A and B meet in
//        |A| cyclic shifts (A_i with B_(i+t) mod |A|), then A and B recurse
//        side by side.  For n = 2^k this is their BR ordering; a half of odd
//        size h instead plays a round robin with byes (round r pairs
//        i + j = 2r mod h and leaves r out) while the two byes A_r, B_r pair
//        across, and the remaining h - 1 shifts t = 1..h-1 come first;
//   id 4 (PJS_MM, n steps): the modulus ordering (mod_pair in pjs.hpp);
//   id 6 (n steps, not in the reference): the block modulus ordering
//        (br_pair in pjs.hpp), the batched kernel's default.
// Every pair meets exactly once per sweep under id 2 (the n/2 pairs
// {i, i + n/2} twice under id 4, the n/2 pairs {2i, 2i + 1} twice under
// id 6); every row is a permutation of 0..n-1
// (needed for race-free column updates); each pair is written with its
// smaller index first.
#include "pjs.hpp"

#include <stdint.h>
#include <stdlib.h>
#include <vector>

namespace {

inline void put(unsigned *tab, unsigned n, unsigned row, unsigned slot, unsigned a, unsigned b)
{
  unsigned *const r = tab + (size_t)row * n + 2u * slot;
  r[0] = (a < b) ? a : b;
  r[1] = (a < b) ? b : a;
}

// rows row0 .. row0 + L - 2, pair slots slot0 .. slot0 + L/2 - 1, for the
// L (even) indices idx[0..L)
void me_fill(const unsigned *idx, unsigned L, unsigned *tab, unsigned n, unsigned row0, unsigned slot0)
{
  if (L == 2u) {
    put(tab, n, row0, slot0, idx[0], idx[1]);
    return;
  }
  const unsigned h = L >> 1u;
  const unsigned *const A = idx, *const B = idx + h;
  if (!(h & 1u)) {
    for (unsigned t = 0u; t < h; ++t)
      for (unsigned i = 0u; i < h; ++i)
        put(tab, n, row0 + t, slot0 + i, A[i], B[(i + t) % h]);
    me_fill(A, h, tab, n, row0 + h, slot0);
    me_fill(B, h, tab, n, row0 + h, slot0 + (h >> 1u));
  }
  else {
    for (unsigned t = 1u; t < h; ++t)
      for (unsigned i = 0u; i < h; ++i)
        put(tab, n, row0 + t - 1u, slot0 + i, A[i], B[(i + t) % h]);
    for (unsigned r = 0u; r < h; ++r) {
      const unsigned row = row0 + h - 1u + r;
      unsigned slot = slot0;
      put(tab, n, row, slot++, A[r], B[r]);
      for (const unsigned *X : {A, B})
        for (unsigned i = 0u; i < h; ++i) {
          const unsigned j = (2u * r + h - i) % h;
          if (i < j)
            put(tab, n, row, slot++, X[i], X[j]);
        }
    }
  }
}

void mantharam_eberlein(unsigned n, unsigned *tab)
{
  std::vector<unsigned> idx(n);
  for (unsigned i = 0u; i < n; ++i)
    idx[i] = i;
  me_fill(idx.data(), n, tab, n, 0u, 0u);
}

void computed(int id, unsigned n, unsigned *tab)
{
  for (unsigned s = 0u; s < n; ++s)
    for (unsigned k = 0u; k < (n >> 1u); ++k) {
      unsigned *const r = tab + (size_t)s * n + 2u * k;
      vj::step_pair(id, n, s, k, r[0], r[1]);
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
  case 6l:
    s = n;
    break;
  default:
    return (unsigned *)NULL;
  }
  unsigned *const tab = (unsigned *)malloc((size_t)n * s * sizeof(unsigned));
  if (!tab)
    return (unsigned *)NULL;
  if (id == 2l)
    mantharam_eberlein(n, tab);
  else
    computed((int)id, n, tab);
  *stp = s;
  return tab;
}

extern "C" void pjs_free(unsigned *tab) { free(tab); }
