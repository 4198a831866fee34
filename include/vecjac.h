/* vecjac.h - C ABI of libvecjac_hip.so, the MI355X (gfx950) implementation of
 * VecJac's one-sided Jacobi SVD and batched 2x2 Hermitian eigensolver.
 *
 * Two groups of entry points:
 *
 *  1. Drop-in replacements with the reference's exact symbol names,
 *     prototypes and argument conventions (Fortran-style, every scalar by
 *     pointer, column-major, ILP64 integers as in common.h:184-190 under
 *     MKL_ILP64).  Arrays are HOST memory, exactly as in the reference; the
 *     library stages them to the GPU, runs there and copies results back.
 *
 *  2. vjh_* extensions on DEVICE memory (pointers from hipMalloc or a torch
 *     CUDA tensor), enqueued on a caller-supplied hipStream_t passed as
 *     void*; plus the batched SVD and the SVD adjoint, which have no
 *     reference counterpart.
 *
 * Error convention (from the reference): a return >= 0 is success (the sweep
 * count for the SVD drivers), -k flags the k-th argument, and the runtime
 * failures return the same negative line codes as the reference's source.
 * Failures of the HIP runtime itself return values <= -100000; see
 * vjh_last_error().
 */
#ifndef VECJAC_H
#define VECJAC_H

#include <stdint.h>

#define VJH_ABI_VERSION 1

typedef int64_t vj_fint;   /* fint: long long under MKL_ILP64 (common.h:184-186) */
typedef uint64_t vj_fnat;  /* fnat: unsigned long long (common.h:185-186) */

#ifdef __cplusplus
extern "C" {
#endif

/* ---- 1. reference-compatible entry points (host memory) ---------------- */

/* Batched real symmetric 2x2 EVD; replaces dbjac2_ (dbjac2.h:6, dbjac2.c:81).
 * n<0: at receives sines; n>0: tangents.  p[i/8] bit i%8 = (l1[i] < l2[i]). */
vj_fint dbjac2_(const vj_fint *n, const double *a11, const double *a22, const double *a21,
                double *c, double *at, double *l1, double *l2, unsigned *p);
/* Batched Hermitian 2x2 EVD; replaces zbjac2_ (zbjac2.h:6, zbjac2.c:108). */
vj_fint zbjac2_(const vj_fint *n, const double *a11, const double *a22, const double *a21r, const double *a21i,
                double *c, double *cat, double *sat, double *l1, double *l2, unsigned *p);
/* FP32 twins; replace sbjac2_ (sbjac2.h:6) and cbjac2_ (cbjac2.h:6); p[i/16]. */
vj_fint sbjac2_(const vj_fint *n, const float *a11, const float *a22, const float *a21,
                float *c, float *at, float *l1, float *l2, unsigned *p);
vj_fint cbjac2_(const vj_fint *n, const float *a11, const float *a22, const float *a21r, const float *a21i,
                float *c, float *cat, float *sat, float *l1, float *l2, unsigned *p);

/* Parallel Jacobi strategy table; replaces pjs() (pjs.h:18, pjs.c:7).
 * id 2 (ME): n-1 steps; id 4 (MM): n steps; n even.  Returns a malloc'd
 * stp x n table (free with free() or pjs_free()), NULL on bad arguments. */
unsigned *pjs(long id, unsigned n, unsigned *stp);
void pjs_free(unsigned *tab);

/* ---- 2. device-memory extensions ---------------------------------------- */

/* Same computations as ?bjac2_ on device arrays, enqueued on `stream`. */
vj_fint vjh_dbjac2(vj_fint n, const double *a11, const double *a22, const double *a21,
                   double *c, double *at, double *l1, double *l2, unsigned *p, void *stream);
vj_fint vjh_zbjac2(vj_fint n, const double *a11, const double *a22, const double *a21r, const double *a21i,
                   double *c, double *cat, double *sat, double *l1, double *l2, unsigned *p, void *stream);
vj_fint vjh_sbjac2(vj_fint n, const float *a11, const float *a22, const float *a21,
                   float *c, float *at, float *l1, float *l2, unsigned *p, void *stream);
vj_fint vjh_cbjac2(vj_fint n, const float *a11, const float *a22, const float *a21r, const float *a21i,
                   float *c, float *cat, float *sat, float *l1, float *l2, unsigned *p, void *stream);

/* Message of the last failure on the calling thread. */
const char *vjh_last_error(void);
int vjh_abi_version(void);

#ifdef __cplusplus
}
#endif

#endif /* !VECJAC_H */
