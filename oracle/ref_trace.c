/* ref_trace M N Gfile JSfile TraceFile
 *
 * Test harness (not product code): runs the reference's own dvjsvd_, built
 * from its sources with its JTRACE output (_ref/libvjref_jtrace.so), on a
 * matrix file (m x n doubles, column-major) and a step-table file (stp x n
 * unsigned, stp = size / 4n), writing the reference's trace to TraceFile.
 * Prints the return value.  Used to check the GPU library's trace files.
 */
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

extern long long dvjsvd_(const uint64_t *m, const uint64_t *n, double *G, const uint64_t *ldG, double *V,
                         const uint64_t *ldV, double *eS, double *fS, const unsigned *js, const unsigned *stp,
                         const unsigned *swp, double *work, unsigned *iwork);

static void *slurp(const char *fn, size_t *sz)
{
  FILE *f = fopen(fn, "rb");
  if (!f)
    return NULL;
  (void)fseek(f, 0, SEEK_END);
  *sz = (size_t)ftell(f);
  (void)fseek(f, 0, SEEK_SET);
  void *p = aligned_alloc(64, (*sz + 63) & ~(size_t)63);
  if (p && fread(p, 1, *sz, f) != *sz) {
    free(p);
    p = NULL;
  }
  fclose(f);
  return p;
}

int main(int argc, char *argv[])
{
  if (argc != 6)
    return 1;
  uint64_t m = strtoull(argv[1], NULL, 10), n = strtoull(argv[2], NULL, 10);
  size_t gsz = 0, jsz = 0;
  double *G = (double *)slurp(argv[3], &gsz);
  unsigned *js = (unsigned *)slurp(argv[4], &jsz);
  if (!G || !js || gsz != m * n * sizeof(double) || jsz % (n * sizeof(unsigned)))
    return 2;
  unsigned stp = (unsigned)(jsz / (n * sizeof(unsigned))), swp = 999u;
  double *V = aligned_alloc(64, n * n * sizeof(double)), *eS = aligned_alloc(64, ((n * 8 + 63) / 64) * 64);
  double *fS = aligned_alloc(64, ((n * 8 + 63) / 64) * 64), *work = aligned_alloc(64, (8 * n + 64) * sizeof(double));
  unsigned *iw = aligned_alloc(64, 64 + n);
  if (!V || !eS || !fS || !work || !iw || strlen(argv[5]) >= 8 * n * sizeof(double))
    return 3;
  iw[0] = 0u;
  strcpy((char *)work, argv[5]);
  printf("%lld\n", dvjsvd_(&m, &n, G, &m, V, &n, eS, fS, js, &stp, &swp, work, iw));
  return 0;
}
