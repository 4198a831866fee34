/* vj_tzvjsvd J M N BaseName
 *
 * Complex twin of vj_tdvjsvd, with the arguments, files, exit codes and
 * report line of the reference's tzvjsvd (tzvjsvd.c:9-199):
 *   reads   BaseName.G          m x n double complex, column-major;
 *   writes  BaseName.S2 / .S4   n singular values (long double);
 *           BaseName.U2 / .U4, BaseName.V2 / .V4   U and V, double complex;
 *   prints  "BaseName",J,m,n,split s,zvjsvd_ s,sweeps,merge V s,merge U s.
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
  const vj_fnat ld = m, ldV = n;
  double *const G = (double *)vj_alloc(2u * m * n * sizeof(double));
  double *const Gr = (double *)vj_alloc(m * n * sizeof(double));
  double *const Gi = (double *)vj_alloc(m * n * sizeof(double));
  if (!G || !Gr || !Gi)
    return 6;
  if (vj_read_cols(gd, m, n, G, ld, 2u * sizeof(double)) || close(gd))
    return 5;
  long double t = vj_now();
  if (zsplit_(&m, &n, (const vj_zdouble *)G, &ld, Gr, &ld, Gi, &ld) < 0)
    return 6;
  const long double ts = vj_now() - t;
  double *const V = (double *)vj_alloc(2u * n * n * sizeof(double));
  double *const Vr = (double *)vj_alloc(n * n * sizeof(double));
  double *const Vi = (double *)vj_alloc(n * n * sizeof(double));
  if (!V || !Vr || !Vi)
    return 7;
  double *const w = (double *)vj_alloc(12u * n * sizeof(double));
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
  t = vj_now();
  const vj_fint o = zvjsvd_(&m, &n, Gr, &ld, Gi, &ld, Vr, &ldV, Vi, &ldV, eS, fS, js, &stp, &swp, work, iwork);
  const long double tj = vj_now() - t;
  vjh_set_jtrace(0); /* the switch is process-wide: off again once this call is traced */
  free(iwork);
  if (vj_write_sigma(bn, (j == 2) ? "S2" : "S4", n, eS, fS))
    return 10;
  free(w);
  t = vj_now();
  if (zmerge_(&n, &n, Vr, &ldV, Vi, &ldV, (vj_zdouble *)V, &ldV) < 0)
    return 11;
  const long double tv = vj_now() - t;
  const int vd = vj_open(bn, (j == 2) ? "V2" : "V4", 1);
  if ((vd < 0) || vj_write_cols(vd, n, n, V, ldV, 2u * sizeof(double)) || close(vd))
    return 11;
  t = vj_now();
  if (zmerge_(&m, &n, Gr, &ld, Gi, &ld, (vj_zdouble *)G, &ld) < 0)
    return 12;
  const long double tg = vj_now() - t;
  const int ud = vj_open(bn, (j == 2) ? "U2" : "U4", 1);
  if ((ud < 0) || vj_write_cols(ud, m, n, G, ld, 2u * sizeof(double)) || close(ud))
    return 12;
  (void)fprintf(stdout, "\"%s\",%1ld,%4llu,%4llu,%15.9Lf,%15.9Lf,%3lld,%15.9Lf,%15.9Lf\n", bn, j, (unsigned long long)m,
                (unsigned long long)n, ts, tj, (long long)o, tv, tg);
  (void)fflush(stdout);
  free(V); free(Vr); free(Vi);
  free(G); free(Gr); free(Gi);
  free(js);
  return EXIT_SUCCESS;
}
