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

/* Integer types of the reference's own interfaces (fint, fnat; common.h:
 * 180-190): 64-bit as under MKL_ILP64, the reference's build default, in
 * libvecjac_hip.so; 32-bit (LP64: int, unsigned) in libvecjac_hip_lp64.so,
 * the same sources compiled with -DVJ_LP64.  A caller compiles against this
 * header with -DVJ_LP64 exactly when it links the LP64 library. */
#ifdef VJ_LP64
typedef int32_t vj_fint;   /* fint: int (common.h:187) */
typedef uint32_t vj_fnat;  /* fnat: unsigned (common.h:188) */
#else
typedef int64_t vj_fint;   /* fint: long long under MKL_ILP64 (common.h:184) */
typedef uint64_t vj_fnat;  /* fnat: unsigned long long (common.h:185) */
#endif
/* Integer types of this library's own entry points (vjh_*): 64-bit in both
 * builds. */
typedef int64_t vjh_int;
typedef uint64_t vjh_nat;
/* element type of interleaved complex arrays: double complex in C (as the
 * reference's prototypes), one (re, im) pair of doubles seen from C++ */
#ifdef __cplusplus
typedef double vj_zdouble;
typedef float vj_cfloat;
#else
typedef double _Complex vj_zdouble;
typedef float _Complex vj_cfloat;
#endif

#ifdef __cplusplus
extern "C" {
#endif

/* ---- 1. reference-compatible entry points (host memory) ---------------- */

/* Batched real symmetric 2x2 EVD; replaces dbjac2_ (dbjac2.h:6, dbjac2.c:81).
 * n<0: at receives sines; n>0: tangents.  p[i/8] bit i%8 = (l1[i] < l2[i]).
 * Returns 1 on success, as the reference's release (OpenMP) build does
 * (dbjac2.c:113, 132), -1 when |n| is not a multiple of 8 (16 for FP32). */
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
/* The internal forms; replace dbjac2i (dbjac2.h:9, dbjac2.c:178), zbjac2i
 * (zbjac2.h:9, zbjac2.c:222), sbjac2i and cbjac2i: n<0 gives secants in at
 * (cat, sat), n>=0 tangents; a block of 8 (FP32: 16) problems is solved only
 * when the low 8 (16) bits of its word p[i/8] (p[i/16]) are not all zero,
 * and those bits are then replaced by the block's l1 < l2 bits; a skipped
 * block's outputs and word are left as they are.  p is read and written. */
vj_fint dbjac2i(const vj_fint *n, const double *a11, const double *a22, const double *a21,
                double *c, double *at, double *l1, double *l2, unsigned *p);
vj_fint zbjac2i(const vj_fint *n, const double *a11, const double *a22, const double *a21r, const double *a21i,
                double *c, double *cat, double *sat, double *l1, double *l2, unsigned *p);
vj_fint sbjac2i(const vj_fint *n, const float *a11, const float *a22, const float *a21,
                float *c, float *at, float *l1, float *l2, unsigned *p);
vj_fint cbjac2i(const vj_fint *n, const float *a11, const float *a22, const float *a21r, const float *a21i,
                float *c, float *cat, float *sat, float *l1, float *l2, unsigned *p);

/* One-sided Jacobi SVD of the m x n matrix G (m >= n), replacing dvjsvd_
 * (dvjsvd.h:10, dvjsvd.c:40).  On return G holds U (unit columns), V the
 * right singular vectors, and sigma_j = fS[j] * 2^eS[j].  js is the stp x n
 * step table (every row a permutation of 0..n-1, else -9), swp the sweep
 * limit, iwork[0] == 0 means V := I on entry.  work holds the trace file
 * name when the trace is on (vjh_set_jtrace), else it is not used.  Returns
 * the number of sweeps, or a negative code with the reference's meaning. */
vj_fint dvjsvd_(const vj_fnat *m, const vj_fnat *n, double *G, const vj_fnat *ldG, double *V, const vj_fnat *ldV,
                double *eS, double *fS, const unsigned *js, const unsigned *stp, const unsigned *swp, double *work,
                unsigned *iwork);
/* Trace switch (the reference's compile-time JTRACE, dvjsvd.c:70-635): with
 * on != 0 the host entry points ?vjsvd_ write the reference's trace file,
 * named by the string in `work`, in the reference's line format; the
 * per-phase times Tn..Te are 0 (one fused GPU step) and Tr is the GPU time
 * of the sweeps.  Process-wide and sticky, as the compile-time switch is:
 * every later host call, from any thread, reads `work` as a file name until
 * it is switched off again, so a caller switches it on around the calls it
 * traces only.  Off by default. */
void vjh_set_jtrace(int on);
/* Rotation form switch (the reference's compile-time USE_SECANTS,
 * vecdef.h:48-52,97-101, dvjsvd.c:411-415): with on != 0 the plain drivers
 * (?vjsvd_, vjh_?vjsvd, vjh_dvjsvd_noalloc) take secants from the 2x2 EVD
 * (the internal form ?bjac2i with n < 0) and every rotation divides by the
 * secant instead of multiplying by the cosine, as a reference built with
 * -DUSE_SECANTS does.  Process-wide and sticky; off by default (the
 * reference's default build).  The batched SVD and the LAPACK-style drivers
 * always use cosines. */
void vjh_set_secants(int on);
int vjh_get_secants(void);

/* Complex (split re/im) twin; replaces zvjsvd_ (zvjsvd.h:10, zvjsvd.c:40). */
vj_fint zvjsvd_(const vj_fnat *m, const vj_fnat *n, double *Gr, const vj_fnat *ldGr, double *Gi, const vj_fnat *ldGi,
                double *Vr, const vj_fnat *ldVr, double *Vi, const vj_fnat *ldVi, double *eS, double *fS,
                const unsigned *js, const unsigned *stp, const unsigned *swp, double *work, unsigned *iwork);
/* FP32 twins; replace svjsvd_ (svjsvd.h:10) and cvjsvd_ (cvjsvd.h:10). */
vj_fint svjsvd_(const vj_fnat *m, const vj_fnat *n, float *G, const vj_fnat *ldG, float *V, const vj_fnat *ldV,
                float *eS, float *fS, const unsigned *js, const unsigned *stp, const unsigned *swp, float *work,
                unsigned *iwork);
vj_fint cvjsvd_(const vj_fnat *m, const vj_fnat *n, float *Gr, const vj_fnat *ldGr, float *Gi, const vj_fnat *ldGi,
                float *Vr, const vj_fnat *ldVr, float *Vi, const vj_fnat *ldVi, float *eS, float *fS,
                const unsigned *js, const unsigned *stp, const unsigned *swp, float *work, unsigned *iwork);

/* Parallel Jacobi strategy table; replaces pjs() (pjs.h:18, pjs.c:7).
 * id 2 (ME): n-1 steps; id 4 (MM): n steps; n even.  Returns a malloc'd
 * stp x n table (free with free() or pjs_free()), NULL on bad arguments.
 * Extension (no reference counterpart): id 6, the block modulus ordering of
 * the batched SVD, n steps (blocks {2i, 2i+1}; round s meets block I with
 * block s - I mod n/2 in two steps), so that the plain drivers can run the
 * batched kernel's sequence of pairs. */
unsigned *pjs(long id, unsigned n, unsigned *stp);

/* Interleaved -> split complex layout; replaces zsplit_ (zsplit.h:6,
 * zsplit.c:3): A (m x n complex, ld ldA) into Ar, Ai (ld ldAr, ldAi).
 * Returns 0, or -4/-6/-8 for a leading dimension below m. */
vj_fint zsplit_(const vj_fnat *m, const vj_fnat *n, const vj_zdouble *A, const vj_fnat *ldA, double *Ar, const vj_fnat *ldAr,
                double *Ai, const vj_fnat *ldAi);
/* Split -> interleaved; replaces zmerge_ (zmerge.h:6, zmerge.c:3). */
vj_fint zmerge_(const vj_fnat *m, const vj_fnat *n, const double *Ar, const vj_fnat *ldAr, const double *Ai,
                const vj_fnat *ldAi, vj_zdouble *A, const vj_fnat *ldA);
void pjs_free(unsigned *tab);

/* LAPACK xGESVJ-style SVD with a sweep limit; replaces dnssvj_ (dnssvj.f:335,
 * called as in tdgesvj.c:7), snssvj_ (snssvj.f:321), znssvj_ (znssvj.f:349,
 * tzgesvj.c:8) and cnssvj_ (cnssvj.f:349).  xGESVJ's arguments plus NSWEEP:
 * JOBA 'G'/'L'/'U', JOBU 'U'/'C'/'N' (complex also 'F'), JOBV 'V'/'A'/'N'
 * (complex also 'J'), M >= N >= 0, A (LDA x N, interleaved complex), SVA(N),
 * V (LDV x N; MV rows for JOBV = 'A'), WORK(LWORK >= max(M+N, 6)) for the
 * real forms, CWORK(LWORK >= M+N) and RWORK(LRWORK >= max(N, 6)) for the
 * complex ones (LWORK or LRWORK = -1: workspace query).  Same argument checks
 * and INFO codes as the reference; INFO = NSWEEP - 1 when NSWEEP sweeps did
 * not converge.  On exit sigma = WORK(1) * SVA(1:N), non-increasing, A holds
 * U (first WORK(3) columns normalised; U * SVA for JOBU = 'N'), V the right
 * singular vectors (or V * (rotations) for 'A'), WORK(2) the number of
 * nonzero values, WORK(3) those above the underflow threshold, WORK(4) the
 * sweep index at exit, WORK(5) max |cos| between the columns paired in the
 * last sweep, WORK(6) max |sin| of its rotations (RWORK for complex).  The
 * SVD runs on the GPU with this library's parallel Jacobi (strategy pjs id 2),
 * so sweep counts and the vectors of clustered values are not LAPACK's.
 * NSWEEP < 1 is rejected (INFO = -14, complex -16; the reference runs no
 * sweep there); INFO below -16 is a failure of the GPU run (vjh_last_error). */
void dnssvj_(const char *joba, const char *jobu, const char *jobv, const vj_fint *m, const vj_fint *n, double *A,
             const vj_fint *lda, double *sva, const vj_fint *mv, double *V, const vj_fint *ldv, double *work,
             const vj_fint *lwork, const vj_fint *nsweep, vj_fint *info);
void snssvj_(const char *joba, const char *jobu, const char *jobv, const vj_fint *m, const vj_fint *n, float *A,
             const vj_fint *lda, float *sva, const vj_fint *mv, float *V, const vj_fint *ldv, float *work,
             const vj_fint *lwork, const vj_fint *nsweep, vj_fint *info);
void znssvj_(const char *joba, const char *jobu, const char *jobv, const vj_fint *m, const vj_fint *n, vj_zdouble *A,
             const vj_fint *lda, double *sva, const vj_fint *mv, vj_zdouble *V, const vj_fint *ldv, vj_zdouble *cwork,
             const vj_fint *lwork, double *rwork, const vj_fint *lrwork, const vj_fint *nsweep, vj_fint *info);
void cnssvj_(const char *joba, const char *jobu, const char *jobv, const vj_fint *m, const vj_fint *n, vj_cfloat *A,
             const vj_fint *lda, float *sva, const vj_fint *mv, vj_cfloat *V, const vj_fint *ldv, vj_cfloat *cwork,
             const vj_fint *lwork, float *rwork, const vj_fint *lrwork, const vj_fint *nsweep, vj_fint *info);

/* ---- 2. device-memory extensions ---------------------------------------- */

/* Same computations as ?bjac2_ (and, vjh_?bjac2i, as ?bjac2i) on device
 * arrays, enqueued on `stream`; they return 0 on success. */
vjh_int vjh_dbjac2(vjh_int n, const double *a11, const double *a22, const double *a21,
                   double *c, double *at, double *l1, double *l2, unsigned *p, void *stream);
vjh_int vjh_zbjac2(vjh_int n, const double *a11, const double *a22, const double *a21r, const double *a21i,
                   double *c, double *cat, double *sat, double *l1, double *l2, unsigned *p, void *stream);
vjh_int vjh_sbjac2(vjh_int n, const float *a11, const float *a22, const float *a21,
                   float *c, float *at, float *l1, float *l2, unsigned *p, void *stream);
vjh_int vjh_cbjac2(vjh_int n, const float *a11, const float *a22, const float *a21r, const float *a21i,
                   float *c, float *cat, float *sat, float *l1, float *l2, unsigned *p, void *stream);
vjh_int vjh_dbjac2i(vjh_int n, const double *a11, const double *a22, const double *a21,
                    double *c, double *at, double *l1, double *l2, unsigned *p, void *stream);
vjh_int vjh_zbjac2i(vjh_int n, const double *a11, const double *a22, const double *a21r, const double *a21i,
                    double *c, double *cat, double *sat, double *l1, double *l2, unsigned *p, void *stream);
vjh_int vjh_sbjac2i(vjh_int n, const float *a11, const float *a22, const float *a21,
                    float *c, float *at, float *l1, float *l2, unsigned *p, void *stream);
vjh_int vjh_cbjac2i(vjh_int n, const float *a11, const float *a22, const float *a21r, const float *a21i,
                    float *c, float *cat, float *sat, float *l1, float *l2, unsigned *p, void *stream);

/* Plain SVD on device memory: same driver as ?vjsvd_, with relaxed shape
 * rules (any even n <= m, any ld >= m, no alignment rule) and a device step
 * table js (validated on the device, -9 / -13 when a row is not a
 * permutation).  The complex forms take one leading dimension for both
 * planes.  vpreset != 0 is the reference's iwork[0] != 0.  stats, if not
 * NULL, receives 66 words: sweeps, sweeps launched, and the number of
 * transformations in each of the first 64 sweeps.  With vpreset == 0 the
 * V updates of each sweep are applied in one pass after it, from per-pair
 * decision records (stp * n/2 of them, each 2 values of the precision,
 * 3 for complex, and a 4-byte code), allocated in stream order with the
 * call; the results are bit for bit those of updating V in every step.
 * Synchronises `stream`. */
vjh_int vjh_dvjsvd(vjh_nat m, vjh_nat n, double *G, vjh_nat ldG, double *V, vjh_nat ldV, double *eS, double *fS,
                   const unsigned *js, unsigned stp, unsigned swp, unsigned vpreset, void *stream, unsigned *stats);
vjh_int vjh_zvjsvd(vjh_nat m, vjh_nat n, double *Gr, double *Gi, vjh_nat ldG, double *Vr, double *Vi, vjh_nat ldV,
                   double *eS, double *fS, const unsigned *js, unsigned stp, unsigned swp, unsigned vpreset,
                   void *stream, unsigned *stats);
vjh_int vjh_svjsvd(vjh_nat m, vjh_nat n, float *G, vjh_nat ldG, float *V, vjh_nat ldV, float *eS, float *fS,
                   const unsigned *js, unsigned stp, unsigned swp, unsigned vpreset, void *stream, unsigned *stats);
vjh_int vjh_cvjsvd(vjh_nat m, vjh_nat n, float *Gr, float *Gi, vjh_nat ldG, float *Vr, float *Vi, vjh_nat ldV,
                   float *eS, float *fS, const unsigned *js, unsigned stp, unsigned swp, unsigned vpreset,
                   void *stream, unsigned *stats);

/* Allocation-free twin of vjh_dvjsvd (no counterpart in the reference), for
 * callers that cannot allocate device or pinned host memory: code that runs
 * between graph captures, tight serving loops, harnesses that forbid
 * hipMalloc.  Same arguments, checks, results (bit for bit) and return codes
 * as vjh_dvjsvd; it calls no hipMalloc/hipFree/hipHostMalloc and creates no
 * event.  Its scratch (control block, table-check flag, per-pair slots) is a
 * static device array of about 0.6 MiB per GPU, held until the process ends,
 * which limits n to VJH_NOALLOC_MAX_N (larger n: -2 with a message).  Every
 * call on one GPU shares that array and leaves its contents behind (each call
 * clears what must start at zero); calls on one GPU are serialised by a
 * per-GPU lock, since overlapping calls would race on the array.  It reads
 * the convergence flag synchronously after each sweep (vjh_dvjsvd queues the
 * sweeps one ahead), so it may be marginally slower.  Synchronises `stream`. */
#define VJH_NOALLOC_MAX_N 16384
vjh_int vjh_dvjsvd_noalloc(vjh_nat m, vjh_nat n, double *G, vjh_nat ldG, double *V, vjh_nat ldV, double *eS,
                           double *fS, const unsigned *js, unsigned stp, unsigned swp, unsigned vpreset, void *stream,
                           unsigned *stats);

/* ---- 3. batched SVD and SVD adjoint (no reference counterpart) --------- */

/* Plan for a batch of nb independent SVDs solved in three persistent
 * launches, one per class: matrices with m above the split (128, complex
 * 64, by default), in the wave form; then the other ("short") ones in the
 * lane-group form, first those whose G exceeds the LDS budget (48 KiB by
 * default), on global memory, then those within it, with G resident in LDS
 * (V in device memory).
 * kind: 0 = FP64 real, 1 = FP64 complex (split re/im), 2 = FP32 real,
 * 3 = FP32 complex.  Host descriptors: m[b], n[b] (n even, 2 <= n <= 128,
 * n <= m <= 1024), element offsets offG[b] of G_b (ld m[b]), offV[b] of V_b (ld n[b])
 * and offS[b] of eS_b/fS_b in the packed device buffers.  waves: waves per
 * workgroup (1, 2, 4 or 8; 0 picks from the median width).  Returns NULL on
 * failure with *status < 0. */
void *vjh_bplan_create(int kind, vjh_int nb, const vjh_int *m, const vjh_int *n, const vjh_int *offG,
                       const vjh_int *offV, const vjh_int *offS, unsigned swp, int waves, vjh_int *status);
/* One run (the launches of the classes) on device buffers: G (A in, U out; U columns of zero singular
 * values are NaN as in the reference), V (out, starts as I), eS/fS (out),
 * info[b] (sweeps or a negative code), transforms[b] (rotations and
 * Gram-Schmidt steps of matrix b; may be NULL).  Gi/Vi are NULL for real
 * kinds.  Asynchronous on `stream`; graph-capturable. */
vjh_int vjh_bplan_run(void *plan, void *G, void *Gi, void *V, void *Vi, void *eS, void *fS, vjh_int *info,
                      uint64_t *transforms, void *stream);
/* Waves per workgroup of the first non-empty launch. */
int vjh_bplan_waves(const void *plan);
/* The plan's choices (out[12]): matrices per class (wave form, lane-group
 * form on global memory, lane-group form in LDS), waves per workgroup of
 * each launch, lanes per pair of each launch (64 for the wave form), the
 * split row count, the LDS budget and the dynamic LDS bytes per workgroup of
 * the LDS launch. */
vjh_int vjh_bplan_config(const void *plan, vjh_int *out);
/* Sets the plan's choices: the split row count (at most 128, complex 64),
 * the LDS budget in bytes (at most 152 KiB; 0 keeps every matrix on global
 * memory) and, where not NULL, waves[3] per workgroup of the three launches
 * (1, 2, 4 or 8) and lanes[3] per pair of the two lane-group launches
 * (lanes[1], lanes[2]: 16, 8 or 4; 8 lanes with at most 4 waves, 4 lanes
 * with at most 2 waves and real rows up to 64).  Returns -5 for a shape that
 * is not built.  The results do not depend on any of them. */
vjh_int vjh_bplan_set(void *plan, int split_rows, vjh_int lds_budget, const int *waves, const int *lanes);
/* Pair evaluations of each matrix in the plan's last run (uint64_t evals[nb]).
 * A pair whose columns did not change since its previous evaluation is not
 * evaluated again, so this can be below sweeps x steps x n/2; it is what the
 * run read.  Synchronises `stream`. */
vjh_int vjh_bplan_evals(const void *plan, uint64_t *evals, void *stream);
/* Elements of G and V (both planes of complex data) that each matrix's
 * solve in the plan's last run read from or wrote to device memory
 * (uint64_t elems[nb]): the column loads and stores of every step, the
 * set-up scan, V := I, the initial scaling and the epilogue; for a matrix
 * whose G is held in LDS, G read once and U written once, and V's columns
 * as above.  Synchronises `stream`. */
vjh_int vjh_bplan_traffic(const void *plan, uint64_t *elems, void *stream);
/* Ordering of the pairs in the plan's runs: 6 (default), the block modulus
 * ordering of pjs(6, n), whose long-column launch keeps the four columns of
 * a block pair in registers for the two steps of a round; or 4, the modulus
 * ordering of pjs(4, n).  A run equals the plain driver run with the same
 * table.  id 0 returns the current ordering; another id returns -2. */
vjh_int vjh_bplan_order(void *plan, int id);
/* Cost model of the plan's processing order (each launch takes its matrices
 * from a queue sorted by decreasing estimated cost, so that the costliest
 * start first): 0, m n^2 (default); 1, n^2 (8 + m/64); 2, n^2 (16 + m/64).
 * model < 0 returns the current model; > 2 returns -2.  The results do not
 * depend on it; the autotuner picks it (step 3). */
vjh_int vjh_bplan_queue(void *plan, int model);
/* Preset V (the batched form of iwork[0] != 0, dvjsvd.c:86-96,143-157): with
 * on = 1 every run takes each V_b (n_b x n_b at offV[b]) as the starting
 * matrix instead of forming I, tracks its max |.| and rescales it as the
 * plain drivers do, and a matrix with a non-finite V fails with -92
 * (complex -104), a zero V with -13 (-17).  Each matrix equals the plain
 * driver run with iwork[0] = 1 on the same table.  Runs with a preset V take
 * the step-by-step form (no block rounds).  on < 0 returns the setting. */
vjh_int vjh_bplan_vpreset(void *plan, int on);
/* Autotuner of a plan, on copies of the input G (Gi) made into the scratch
 * buffer Gw (Gwi), each launch timed on its own with HIP events on `stream`:
 * (1) for the split row count and half of it, each launch at 1, 2, 4 and 8
 * waves per workgroup (counts that would spill registers are skipped) and
 * the lane-group launches also at 8 and 4 lanes per pair, keeping the split
 * and shapes of the fastest total; (2) for LDS budgets of 0, 24, 48, 96 and
 * 152 KiB, the two short launches at each shape, keeping the fastest
 * budget and shapes; (3) for each queue cost model (vjh_bplan_queue), the
 * whole plan twice, keeping the fastest model.  The choices stay in the
 * plan (vjh_bplan_config, vjh_bplan_queue).  ms[4] (may be NULL) receives the run times of the
 * four counts at the default split, -1 for skipped counts.  V, eS, fS and
 * info are overwritten.  Returns the count of the first non-empty launch,
 * or a negative code.  Synchronises `stream`. */
vjh_int vjh_bplan_autotune(void *plan, const void *G, const void *Gi, void *Gw, void *Gwi, void *V, void *Vi, void *eS,
                           void *fS, vjh_int *info, float *ms, void *stream);
void vjh_bplan_destroy(void *plan);
/* zsplit_ / zmerge_ on device memory (interleaved A as pairs of doubles),
 * one coalesced pass on `stream`, no host round trip. */
vjh_int vjh_zsplit(vjh_nat m, vjh_nat n, const double *A, vjh_nat ldA, double *Ar, vjh_nat ldAr, double *Ai, vjh_nat ldAi,
                   void *stream);
vjh_int vjh_zmerge(vjh_nat m, vjh_nat n, const double *Ar, vjh_nat ldAr, const double *Ai, vjh_nat ldAi, double *A, vjh_nat ldA,
                   void *stream);

/* One-shot FP64 batched SVD (plan, run, synchronise, destroy). */
vjh_int vjh_dbvjsvd(vjh_int nb, const vjh_int *m, const vjh_int *n, const vjh_int *offG, const vjh_int *offV,
                    const vjh_int *offS, double *G, double *V, double *eS, double *fS, vjh_int *info,
                    uint64_t *transforms, unsigned swp, int waves, void *stream);

/* Adjoint of the thin SVD A = U diag(s) V^T for nb matrices of m x n (any
 * 1 <= n <= m; distinct nonzero s): given U (m x n), s, V (n x n) and the
 * cotangents dU, ds, dV, writes
 *   dA = U [diag(ds) + (F o (U^T dU - dU^T U)) S + S (F o (V^T dV - dV^T V))] V^T
 *        + (I - U U^T) dU S^-1 V^T,   F_ij = 1 / (s_j^2 - s_i^2), F_ii = 0.
 * Matrix b: U, dU, dA at b*m*n (ld m); V, dV at b*n*n (ld n); s, ds at b*n.
 * The n x n intermediates stay in LDS up to n = 96 (FP64; complex 64);
 * wider matrices take a stream-ordered workspace (hipMallocAsync).
 * Asynchronous on `stream`. */
vjh_int vjh_dsvd_bwd(vjh_int nb, vjh_int m, vjh_int n, const double *U, const double *s, const double *V,
                     const double *dU, const double *ds, const double *dV, double *dA, void *stream);
vjh_int vjh_ssvd_bwd(vjh_int nb, vjh_int m, vjh_int n, const float *U, const float *s, const float *V,
                     const float *dU, const float *ds, const float *dV, float *dA, void *stream);
/* Complex form (split real and imaginary planes, A = U diag(s) V^H), for a
 * loss that does not depend on the phases of the singular vector pairs
 * (u_k, v_k) -> (u_k e^{i phi}, v_k e^{i phi}): with J = U^H dU, K = V^H dV,
 *   dA = U [diag(ds) + (F o (J - J^H)) S + S (F o (K - K^H))
 *           + i diag(Im J_kk - Im K_kk) / (2 S)] V^H + (I - U U^H) dU S^-1 V^H.
 * The cotangents follow the convention dL = Re sum(conj(dX) o dX') (torch's
 * complex gradients). */
vjh_int vjh_zsvd_bwd(vjh_int nb, vjh_int m, vjh_int n, const double *Ur, const double *Ui, const double *s,
                     const double *Vr, const double *Vi, const double *dUr, const double *dUi, const double *ds,
                     const double *dVr, const double *dVi, double *dAr, double *dAi, void *stream);
vjh_int vjh_csvd_bwd(vjh_int nb, vjh_int m, vjh_int n, const float *Ur, const float *Ui, const float *s,
                     const float *Vr, const float *Vi, const float *dUr, const float *dUi, const float *ds,
                     const float *dVr, const float *dVi, float *dAr, float *dAi, void *stream);

/* Message of the last failure on the calling thread. */
const char *vjh_last_error(void);
int vjh_abi_version(void);
/* Bits of vj_fint in the loaded library: 64 (libvecjac_hip.so) or 32
 * (libvecjac_hip_lp64.so). */
int vjh_fint_bits(void);

#ifdef __cplusplus
}
#endif

#endif /* !VECJAC_H */
