/* Oracle (test infrastructure only): instantiates the restated kernels and
 * drivers for the four precisions of the reference (d, z, s, c). */
#include "vjo.h"
#include "vjo_prim.h"

#include <stdlib.h>
#ifdef __x86_64__
#include <xmmintrin.h>
#endif

/* the reference refuses to run unless MXCSR holds its default (vec.h:19-25) */
static inline int vjo_fpenv_bad(void)
{
#ifdef __x86_64__
  return ((_mm_getcsr() & 0xFFC0u) != 0x1F80u);
#else
  return 0;
#endif
}
#define VJO_MISALIGNED(x) (((uintptr_t)(x)) & 0x3Fu)

/* 0: column norms as the reference's long-double dnormf_ and dots as its
 * ddpscl_/zdpscl_ (bit-identical to the reference build); 1: the GPU's
 * correctly rounded norms and dots */
static int vjo_norm_mode = 0;
void vjo_set_norm_mode(int mode) { vjo_norm_mode = mode; }
int vjo_get_norm_mode(void) { return vjo_norm_mode; }

/* 0: the rotations scale by the cosine (the reference's default build);
 * 1: the reference built with USE_SECANTS (vecdef.h:48-52,97-101): the
 * drivers ask the 2x2 EVD for secants (dvjsvd.c:411-415, dbjac2i with
 * n < 0) and the rotations divide by them */
static int vjo_secants = 0;
void vjo_set_secants(int on) { vjo_secants = on; }
int vjo_get_secants(void) { return vjo_secants; }
#define VJO_CSCALE(v, c) (vjo_secants ? ((v) / (c)) : ((v) * (c)))

/* ------------------------------ double ------------------------------ */
#define T double
#define SFX(x) x##_d
#define RN(x) vjo_d##x
#define CN(x) vjo_z##x
#define RED(v) red8_d(v)
#define VL 8u
#define BIG_EXP 1020.0
#define SQRT_HUGE 1.34078079299425956E+154
#define TMAX DBL_MAX
#define TMIN_SUB DBL_TRUE_MIN
#define TEPS DBL_EPSILON
#define FIN_EXP 1023.0
#define ROT_EXP_R 1022
#define ROT_EXP_C 1021
#define FMA fma
#define SQRT sqrt
#define FMAX fmax
#define FMIN fmin
#define SCALBN scalbn
#define COPYSIGN copysign
#include "vjo_jac2_tmpl.h"
#include "vjo_l1_tmpl.h"
#define CPLX 0
#include "vjo_drv_tmpl.h"
#undef CPLX
#define CPLX 1
#include "vjo_drv_tmpl.h"
#undef CPLX
#undef T
#undef SFX
#undef RN
#undef CN
#undef RED
#undef VL
#undef BIG_EXP
#undef SQRT_HUGE
#undef TMAX
#undef TMIN_SUB
#undef TEPS
#undef FIN_EXP
#undef ROT_EXP_R
#undef ROT_EXP_C
#undef FMA
#undef SQRT
#undef FMAX
#undef FMIN
#undef SCALBN
#undef COPYSIGN

/* ------------------------------ float ------------------------------ */
#define T float
#define SFX(x) x##_f
#define RN(x) vjo_s##x
#define CN(x) vjo_c##x
#define RED(v) red16_f(v)
#define VL 16u
#define BIG_EXP 124.0f
#define SQRT_HUGE 1.844674297e+19f
#define TMAX FLT_MAX
#define TMIN_SUB FLT_TRUE_MIN
#define TEPS FLT_EPSILON
#define FIN_EXP 127.0f
#define ROT_EXP_R 126
#define ROT_EXP_C 125
#define FMA fmaf
#define SQRT sqrtf
#define FMAX fmaxf
#define FMIN fminf
#define SCALBN scalbnf
#define COPYSIGN copysignf
#include "vjo_jac2_tmpl.h"
#include "vjo_l1_tmpl.h"
#define CPLX 0
#include "vjo_drv_tmpl.h"
#undef CPLX
#define CPLX 1
#include "vjo_drv_tmpl.h"
#undef CPLX

/* ------------------- batches of independent SVDs ------------------- */
#define VJO_BATCH(NAME, CORE, TT)                                                        \
int64_t NAME(int64_t nb, const int64_t *m, const int64_t *n, TT *G, const int64_t *offG,  \
             TT *V, const int64_t *offV, TT *eS, TT *fS, const int64_t *offS,             \
             const unsigned *const *js_by_n, const unsigned *stp_by_n, unsigned swp,      \
             int64_t *ret)                                                                \
{                                                                                         \
  int64_t bad = 0;                                                                        \
  _Pragma("omp parallel for schedule(dynamic,1) reduction(+:bad)")                        \
  for (int64_t b = 0; b < nb; ++b) {                                                      \
    const vjo_fnat mm = (vjo_fnat)m[b], nn = (vjo_fnat)n[b];                              \
    unsigned iw = 0u;                                                                     \
    const vjo_fint o = CORE(0, &mm, &nn, G + offG[b], &mm, (TT*)NULL, &mm, V + offV[b],   \
                            &nn, (TT*)NULL, &nn, eS + offS[b], fS + offS[b],              \
                            js_by_n[nn], stp_by_n + nn, &swp, (TT*)NULL, &iw,             \
                            (vjo_stats*)NULL);                                            \
    if (ret) ret[b] = (int64_t)o;                                                         \
    if (o < 0) ++bad;                                                                     \
  }                                                                                       \
  return bad;                                                                             \
}
VJO_BATCH(vjo_dvjsvd_batch, vjo_dvjsvd_core, double)
VJO_BATCH(vjo_svjsvd_batch, vjo_svjsvd_core, float)
