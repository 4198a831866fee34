/* Oracle (test infrastructure only): CPU restatement of the reference's
 * one-sided Jacobi SVD and batched 2x2 Hermitian EVD.  Only tests/, the
 * smoke check and the benchmark's CPU-baseline leg may load liboracle_vj.so,
 * and only to check results, never as the measured or shipped path.
 *
 * Names mirror the reference with a vjo_ prefix and no trailing underscore;
 * arguments are passed the way the reference passes them (by pointer for the
 * scalars of the Fortran-style entry points). */
#ifndef VJO_H
#define VJO_H

#include <stddef.h>
#include <stdint.h>

typedef int64_t vjo_fint;   /* fint under MKL_ILP64 (common.h:184-190) */
typedef uint64_t vjo_fnat;

#define VJO_MAX_TRACE 64u
typedef struct {
  unsigned sweeps;            /* return value on success */
  int sT, sV;                 /* total G and V power-of-two shifts */
  unsigned rescales;          /* in-loop rescales of G */
  uint64_t swt[VJO_MAX_TRACE]; /* transformations per sweep (dvjsvd.c:596) */
} vjo_stats;

enum { VJO_TAN = 0, VJO_SIN = 1, VJO_SEC = 2 };

#ifdef __cplusplus
extern "C" {
#endif
/* 2x2 EVD (dbjac2.c, zbjac2.c, sbjac2.c, cbjac2.c) */
vjo_fint vjo_dbjac2(const vjo_fint *n, const double *a11, const double *a22, const double *a21, double *c, double *at, double *l1, double *l2, unsigned *p);
vjo_fint vjo_dbjac2i(const vjo_fint *n, const double *a11, const double *a22, const double *a21, double *c, double *at, double *l1, double *l2, unsigned *p);
vjo_fint vjo_zbjac2(const vjo_fint *n, const double *a11, const double *a22, const double *a21r, const double *a21i, double *c, double *cat, double *sat, double *l1, double *l2, unsigned *p);
vjo_fint vjo_zbjac2i(const vjo_fint *n, const double *a11, const double *a22, const double *a21r, const double *a21i, double *c, double *cat, double *sat, double *l1, double *l2, unsigned *p);
vjo_fint vjo_sbjac2(const vjo_fint *n, const float *a11, const float *a22, const float *a21, float *c, float *at, float *l1, float *l2, unsigned *p);
vjo_fint vjo_sbjac2i(const vjo_fint *n, const float *a11, const float *a22, const float *a21, float *c, float *at, float *l1, float *l2, unsigned *p);
vjo_fint vjo_cbjac2(const vjo_fint *n, const float *a11, const float *a22, const float *a21r, const float *a21i, float *c, float *cat, float *sat, float *l1, float *l2, unsigned *p);
vjo_fint vjo_cbjac2i(const vjo_fint *n, const float *a11, const float *a22, const float *a21r, const float *a21i, float *c, float *cat, float *sat, float *l1, float *l2, unsigned *p);

/* level-1 kernels (double; the float twins carry the s/c prefix) */
double vjo_dnorm2(vjo_fnat m, const double *x, double *e0, double *f0);
double vjo_znorm2(vjo_fnat m, const double *xr, const double *xi, double *e0, double *f0);
double vjo_ddpscl(vjo_fnat m, const double *x, const double *y, const double *e, const double *f);
void vjo_zdpscl(vjo_fnat m, const double *xr, const double *xi, const double *yr, const double *yi, const double *e, const double *f, double *zr, double *zi);
double vjo_djrot(vjo_fint n, double *x, double *y, double c, double at);
double vjo_djrotf(vjo_fint n, double *x, double *y, double c, double at);
double vjo_zjrot(vjo_fint n, double *xr, double *xi, double *yr, double *yi, double c, double cat, double sat);
double vjo_zjrotf(vjo_fint n, double *xr, double *xi, double *yr, double *yi, double c, double cat, double sat);
double vjo_dgsscl(vjo_fint m, double t, double *x, double *y, const double *e, const double *f);
double vjo_zgsscl(vjo_fint m, double tr, double ti, double *xr, double *xi, double *yr, double *yi, const double *e, const double *f);
void vjo_dscale(vjo_fnat m, vjo_fnat n, double *A, vjo_fnat ldA, vjo_fint e);
double vjo_dnormx(vjo_fnat m, vjo_fnat n, const double *A, vjo_fnat ldA);
float vjo_snorm2(vjo_fnat m, const float *x, float *e0, float *f0);
float vjo_sdpscl(vjo_fnat m, const float *x, const float *y, const float *e, const float *f);
void vjo_cdpscl(vjo_fnat m, const float *xr, const float *xi, const float *yr, const float *yi, const float *e, const float *f, float *zr, float *zi);
float vjo_sjrot(vjo_fint n, float *x, float *y, float c, float at);
float vjo_sgsscl(vjo_fint m, float t, float *x, float *y, const float *e, const float *f);

/* column-norm definition used by the drivers: 0 = the reference's long-double
 * dnormf_ (default), 1 = the GPU's correctly rounded scaled norm */
void vjo_set_norm_mode(int mode);
int vjo_get_norm_mode(void);
/* rotation form used by the drivers: 0 = cosines (default), 1 = secants (the
 * reference's USE_SECANTS build, vecdef.h:97-101, dvjsvd.c:411-415) */
void vjo_set_secants(int on);
int vjo_get_secants(void);

/* SVD drivers: reference semantics (dvjsvd.h:10, zvjsvd.h:10, svjsvd.h:10, cvjsvd.h:10) */
vjo_fint vjo_dvjsvd(const vjo_fnat *m, const vjo_fnat *n, double *G, const vjo_fnat *ldG, double *V, const vjo_fnat *ldV, double *eS, double *fS, const unsigned *js, const unsigned *stp, const unsigned *swp, double *work, unsigned *iwork);
vjo_fint vjo_zvjsvd(const vjo_fnat *m, const vjo_fnat *n, double *Gr, const vjo_fnat *ldGr, double *Gi, const vjo_fnat *ldGi, double *Vr, const vjo_fnat *ldVr, double *Vi, const vjo_fnat *ldVi, double *eS, double *fS, const unsigned *js, const unsigned *stp, const unsigned *swp, double *work, unsigned *iwork);
vjo_fint vjo_svjsvd(const vjo_fnat *m, const vjo_fnat *n, float *G, const vjo_fnat *ldG, float *V, const vjo_fnat *ldV, float *eS, float *fS, const unsigned *js, const unsigned *stp, const unsigned *swp, float *work, unsigned *iwork);
vjo_fint vjo_cvjsvd(const vjo_fnat *m, const vjo_fnat *n, float *Gr, const vjo_fnat *ldGr, float *Gi, const vjo_fnat *ldGi, float *Vr, const vjo_fnat *ldVr, float *Vi, const vjo_fnat *ldVi, float *eS, float *fS, const unsigned *js, const unsigned *stp, const unsigned *swp, float *work, unsigned *iwork);

/* relaxed cores: strict=0 drops the multiple-of-VL, alignment and MXCSR
 * checks (any even n, any m >= n); stats may be NULL */
vjo_fint vjo_dvjsvd_core(int strict, const vjo_fnat *m, const vjo_fnat *n, double *G, const vjo_fnat *ldG, double *Gi, const vjo_fnat *ldGi, double *V, const vjo_fnat *ldV, double *Vi, const vjo_fnat *ldVi, double *eS, double *fS, const unsigned *js, const unsigned *stp, const unsigned *swp, double *work, unsigned *iwork, vjo_stats *stats);
vjo_fint vjo_zvjsvd_core(int strict, const vjo_fnat *m, const vjo_fnat *n, double *Gr, const vjo_fnat *ldGr, double *Gi, const vjo_fnat *ldGi, double *Vr, const vjo_fnat *ldVr, double *Vi, const vjo_fnat *ldVi, double *eS, double *fS, const unsigned *js, const unsigned *stp, const unsigned *swp, double *work, unsigned *iwork, vjo_stats *stats);
vjo_fint vjo_svjsvd_core(int strict, const vjo_fnat *m, const vjo_fnat *n, float *G, const vjo_fnat *ldG, float *Gi, const vjo_fnat *ldGi, float *V, const vjo_fnat *ldV, float *Vi, const vjo_fnat *ldVi, float *eS, float *fS, const unsigned *js, const unsigned *stp, const unsigned *swp, float *work, unsigned *iwork, vjo_stats *stats);
vjo_fint vjo_cvjsvd_core(int strict, const vjo_fnat *m, const vjo_fnat *n, float *Gr, const vjo_fnat *ldGr, float *Gi, const vjo_fnat *ldGi, float *Vr, const vjo_fnat *ldVr, float *Vi, const vjo_fnat *ldVi, float *eS, float *fS, const unsigned *js, const unsigned *stp, const unsigned *swp, float *work, unsigned *iwork, vjo_stats *stats);

/* many independent relaxed-core SVDs over a thread pool (OpenMP) for the
 * CPU baseline of the batched workload: matrix b has m[b] x n[b] G at
 * G + offG[b] (ld = m[b]) and n[b] x n[b] V at V + offV[b]; js_by_n[n] is
 * the step table for width n with stp_by_n[n] rows.  Returns the number of
 * matrices whose driver returned a negative code. */
int64_t vjo_dvjsvd_batch(int64_t nb, const int64_t *m, const int64_t *n, double *G, const int64_t *offG, double *V, const int64_t *offV, double *eS, double *fS, const int64_t *offS, const unsigned *const *js_by_n, const unsigned *stp_by_n, unsigned swp, int64_t *ret);
int64_t vjo_svjsvd_batch(int64_t nb, const int64_t *m, const int64_t *n, float *G, const int64_t *offG, float *V, const int64_t *offV, float *eS, float *fS, const int64_t *offS, const unsigned *const *js_by_n, const unsigned *stp_by_n, unsigned swp, int64_t *ret);
#ifdef __cplusplus
}
#endif

#endif /* !VJO_H */
