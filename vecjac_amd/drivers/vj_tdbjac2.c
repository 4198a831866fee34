/* vj_tdbjac2 filename lg(batch_size) #batches
 *
 * Command-line driver with the arguments, input files and CSV report of the
 * reference's tdbjac2 (tdbjac2.c:5-212), running dbjac2_ of libvecjac_hip.so
 * on the GPU.  The inputs are the reference's generated test files:
 *   filename.f / .g / .h   a11, a22, a21 of the 2x2 problems (doubles);
 *   filename.k / .l        the eigenvalues they were generated from;
 * each holding #batches batches of 2^lg problems.  Per batch one line
 *   B,Ts,ORT,REN,RLN,RLX,RLM
 * with the dbjac2_ time (sine form) and, in long double, the largest
 *   ORT  |cos^2 + sin^2 - 1| of the computed rotation,
 *   REN  ||U diag(l1, l2) U^T - A||_F / ||A||_F,
 *   RLN  ||L - l||_2 / ||l||_2 of the sorted |eigenvalues| against the
 *        generating ones, and RLX / RLM its larger / smaller component
 *        relative errors.
 */
#include "vjio.h"

/* Frobenius norm of the symmetric 2x2 [[a, b], [b, d]] */
static long double fro2(long double a, long double d, long double b)
{
  return hypotl(hypotl(a, d), sqrtl(2.0L) * b);
}

int main(int argc, char *argv[])
{
  if (argc != 4) {
    (void)fprintf(stderr, "%s filename lg(batch_size) #batches\n", *argv);
    return EXIT_FAILURE;
  }
  const size_t n = (size_t)1u << strtoull(argv[2], NULL, 10);
  if (n % 8u) {
    (void)fprintf(stderr, "batch_size has to be a multiple of 8.\n");
    return EXIT_FAILURE;
  }
  const size_t b = (size_t)strtoull(argv[3], NULL, 10);
  if (!b)
    return EXIT_SUCCESS;
  const char *ex[5] = {"f", "g", "h", "k", "l"};
  int fd[5];
  for (int k = 0; k < 5; ++k)
    if ((fd[k] = vj_open(argv[1], ex[k], 0)) < 0) {
      (void)fprintf(stderr, "Cannot open %s.%s for reading!\n", argv[1], ex[k]);
      return EXIT_FAILURE;
    }
  double *a11 = (double *)vj_alloc(n * sizeof(double)), *a22 = (double *)vj_alloc(n * sizeof(double));
  double *a21 = (double *)vj_alloc(n * sizeof(double)), *c = (double *)vj_alloc(n * sizeof(double));
  double *sn = (double *)vj_alloc(n * sizeof(double)), *l1 = (double *)vj_alloc(n * sizeof(double));
  double *l2 = (double *)vj_alloc(n * sizeof(double)), *g1 = (double *)vj_alloc(n * sizeof(double));
  double *g2 = (double *)vj_alloc(n * sizeof(double));
  unsigned *p = (unsigned *)vj_alloc((n / 8u) * sizeof(unsigned));
  if (!a11 || !a22 || !a21 || !c || !sn || !l1 || !l2 || !g1 || !g2 || !p)
    return EXIT_FAILURE;
  (void)fprintf(stdout, "\"B\",\"Ts\",\"ORT\",\"REN\",\"RLN\",\"RLX\",\"RLM\"\n");
  (void)fflush(stdout);
  for (size_t j = 0; j < b; ++j) {
    const off_t off = (off_t)(j * n * sizeof(double));
    double *in[5] = {a11, a22, a21, g1, g2};
    for (int k = 0; k < 5; ++k)
      if ((ssize_t)(n * sizeof(double)) != pread(fd[k], in[k], n * sizeof(double), off))
        return EXIT_FAILURE;
    const vj_fint nn = -(vj_fint)n;
    const long double t0 = vj_now();
    const vj_fint rc = dbjac2_(&nn, a11, a22, a21, c, sn, l1, l2, p);
    const long double ts = vj_now() - t0;
    if (rc < 0)
      return EXIT_FAILURE;
    long double ort = 0.0L, ren = 0.0L, rln = 0.0L, rlx = 0.0L, rlm = 0.0L;
    for (size_t i = 0; i < n; ++i) {
      const long double C = c[i], S = sn[i], L1 = l1[i], L2 = l2[i];
      ort = fmaxl(ort, fabsl(fmal(C, C, fmal(S, S, -1.0L))));
      // A = U diag(L1, L2) U^T with U = [[C, -S], [S, C]]
      const long double r11 = fmal(L1, C * C, L2 * S * S) - (long double)a11[i];
      const long double r22 = fmal(L2, C * C, L1 * S * S) - (long double)a22[i];
      const long double r21 = (L1 - L2) * C * S - (long double)a21[i];
      const long double an = fro2(a11[i], a22[i], a21[i]);
      if (an > 0.0L)
        ren = fmaxl(ren, fro2(r11, r22, r21) / an);
      long double X = fabsl(L1), Y = fabsl(L2), x = fabsl((long double)g1[i]), y = fabsl((long double)g2[i]);
      if (X > Y) { const long double t = X; X = Y; Y = t; }
      if (x > y) { const long double t = x; x = y; y = t; }
      const long double ex_ = fabsl(Y - y), en_ = fabsl(X - x);
      if (y > 0.0L)
        rlx = fmaxl(rlx, ex_ / y);
      if (x > 0.0L)
        rlm = fmaxl(rlm, en_ / x);
      const long double ln = hypotl(y, x);
      if (ln > 0.0L)
        rln = fmaxl(rln, hypotl(ex_, en_) / ln);
    }
    (void)fprintf(stdout, "%zu,%15.9Lf,%#.21LE,%#.21LE,%#.21LE,%#.21LE,%#.21LE\n", j, ts, ort, ren, rln, rlx, rlm);
    (void)fflush(stdout);
  }
  for (int k = 0; k < 5; ++k)
    (void)close(fd[k]);
  free(a11); free(a22); free(a21); free(c); free(sn); free(l1); free(l2); free(g1); free(g2); free(p);
  return EXIT_SUCCESS;
}
