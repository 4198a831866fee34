/* vj_tdvjsvd J M N BaseName
 *
 * Command-line driver with the arguments, files, exit codes and report line
 * of the reference's tdvjsvd (tdvjsvd.c:7-168), running dvjsvd_ of
 * libvecjac_hip.so on the GPU:
 *   reads   BaseName.G          m x n doubles, column-major;
 *   writes  BaseName.S2 / .S4   n singular values (long double), J = 2 / 4;
 *           BaseName.U2 / .U4   U, m x n;  BaseName.V2 / .V4  V, n x n;
 *   prints  "BaseName",J,m,n,seconds,sweeps  (the dvjsvd_ return value).
 * With the environment variable VJ_JTRACE set, the trace of the reference's
 * JTRACE build is written to BaseName.J (see vjh_set_jtrace).
 * J is the pjs strategy: 2 (ME) or 4 (MM).  n must be even with n/2 a
 * multiple of 8 and n <= m, as the reference requires.
 */
#include "vjio.h"

int main(int argc, char *argv[])
{
  if (argc != 5) {
    (void)fprintf(stderr, "%s J M N BaseName\n", *argv);
    return 1;
  }
  const vj_fnat m = (vj_fnat)strtoull(argv[2], NULL, 10);
  if (!m)
    return 3;
  const vj_fnat n = (vj_fnat)strtoull(argv[3], NULL, 10);
  if (!n || (n > m) || (n & 1u) || ((n >> 1u) & 7u))
    return 4;
  const long j = atol(argv[1]);
  if ((j != 2) && (j != 4))
    return 2;
  unsigned stp = 0u;
  unsigned *const js = pjs(j, (unsigned)n, &stp);
  if (!js)
    return 2;
  const char *const bn = argv[4];
  if (!*bn)
    return 5;
  const int gd = vj_open(bn, "G", 0);
  if (gd < 0)
    return 5;
  const vj_fnat ldG = m, ldV = n;
  double *const G = (double *)vj_alloc(m * n * sizeof(double));
  if (!G)
    return 6;
  if (vj_read_cols(gd, m, n, G, ldG, sizeof(double)) || close(gd))
    return 5;
  double *const V = (double *)vj_alloc(n * n * sizeof(double));
  if (!V)
    return 7;
  double *const w = (double *)vj_alloc(10u * n * sizeof(double));
  if (!w)
    return 8;
  double *const eS = w, *const fS = w + n, *const work = w + 2u * n;
  unsigned *const iwork = (unsigned *)vj_alloc(((n >> 3u) + 1u) * sizeof(unsigned));
  if (!iwork)
    return 9;
  *iwork = 0u;
  if (getenv("VJ_JTRACE")) { /* the trace of the reference's JTRACE build: BaseName.J */
    (void)snprintf((char *)work, 8u * n * sizeof(double), "%s.%ld", bn, j);
    vjh_set_jtrace(1);
  }
  const unsigned swp = 999u;
  const long double t0 = vj_now();
  const vj_fint o = dvjsvd_(&m, &n, G, &ldG, V, &ldV, eS, fS, js, &stp, &swp, work, iwork);
  const long double t1 = vj_now();
  vjh_set_jtrace(0); /* the switch is process-wide: off again once this call is traced */
  (void)fprintf(stdout, "\"%s\",%1ld,%4llu,%4llu,%15.9Lf,%3lld\n", bn, j, (unsigned long long)m, (unsigned long long)n,
                t1 - t0, (long long)o);
  (void)fflush(stdout);
  free(iwork);
  if (vj_write_sigma(bn, (j == 2) ? "S2" : "S4", n, eS, fS))
    return 10;
  free(w);
  const int vd = vj_open(bn, (j == 2) ? "V2" : "V4", 1);
  if ((vd < 0) || vj_write_cols(vd, n, n, V, ldV, sizeof(double)) || close(vd))
    return 11;
  free(V);
  const int ud = vj_open(bn, (j == 2) ? "U2" : "U4", 1);
  if ((ud < 0) || vj_write_cols(ud, m, n, G, ldG, sizeof(double)) || close(ud))
    return 12;
  free(G);
  free(js);
  return EXIT_SUCCESS;
}
