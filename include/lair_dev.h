/*
 * lair_dev.h -- C ABI of the MI355X (gfx950) LU factor-and-solve library.
 *
 * Drop-in boundary for the LU path of vinesystems/lair 0.8.0.  Each entry point
 * names the reference interface it replaces.  Conventions shared by all calls:
 *
 *  * Matrices are described like an ndarray view: a base pointer and element
 *    strides `rs` (between rows) and `cs` (between columns), which may be
 *    negative.  The standard (C-order) layout is rs = ncols, cs = 1.
 *  * Pivots are 0-based absolute row indices in int64 (the reference's
 *    Vec<usize>): at step i, row i was interchanged with row ipiv[i].  This is
 *    LAPACK's ipiv minus one.
 *  * `singular` receives the reference's Option<usize>: -1 for None, otherwise
 *    the LAST column whose pivot was exactly zero (src/lapack/getrf.rs:72-73).
 *  * Host-pointer calls are blocking and thread-safe.  `_device` calls take
 *    device pointers obtained from lair_malloc() and run on the context's
 *    stream; they return after the work is queued unless stated otherwise.
 *
 * Status codes: 0 success; negative -i means argument i was invalid (LAPACK
 * style); positive values are the LAIR_ERR_* codes below.
 */
#ifndef LAIR_DEV_H
#define LAIR_DEV_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef int32_t lair_status;

#define LAIR_OK 0
#define LAIR_ERR_SHAPE 1    /* InvalidInput::Shape (src/lib.rs:26-27) */
#define LAIR_ERR_SINGULAR 2 /* InvalidInput::Value("`a` is a singular matrix") (src/equation.rs:56) */
#define LAIR_ERR_DEVICE 3   /* HIP runtime or launch failure (no reference analogue) */
#define LAIR_ERR_NOMEM 4    /* device allocation failed (no reference analogue) */
#define LAIR_ERR_NOCONV 5   /* refinement did not reach the tolerance; FP64 fallback was used */
#define LAIR_ERR_SYNC 6     /* a cooperative kernel's bounded wait expired (no reference analogue) */
#define LAIR_ERR_UNSUPPORTED 7

typedef struct lair_context lair_context;

/* Precision of the O(n^3) factorization in the mixed solver. */
#define LAIR_PREC_FP64 0
#define LAIR_PREC_FP32 1  /* FP32 storage, FP32 MFMA (v_mfma_f32_16x16x4_f32) trailing update */
#define LAIR_PREC_BF16X6 2 /* FP32 storage, trailing update as six BF16 MFMA products
                              (3-way bf16 split of each operand), FP32 accumulation */
#define LAIR_PREC_BF16 3   /* FP32 storage, single BF16 MFMA product (diagnostic; does not
                              converge on ill-conditioned systems) */
/* Flag OR'd into the precision of the mixed solver: choose each 16-column
 * leaf's pivot rows by tournament pivoting (a reduction tree of small
 * partial-pivoting eliminations, one cross-workgroup exchange per leaf)
 * instead of one global argmax per column.  The pivot order then differs from
 * partial pivoting; the refined FP64 answer meets the same tolerance. */
#define LAIR_PIVOT_TOURNAMENT 0x100

/* ------------------------------------------------------------ context --- */

/* Creates a context on the current HIP device: a stream, workspace and the
 * autotuner cache.  The host-pointer calls below use a lazily created,
 * mutex-guarded default context when `ctx` is NULL. */
lair_status lair_context_create(lair_context **ctx);
lair_status lair_context_destroy(lair_context *ctx);
/* Blocks until all work queued on the context's stream has finished, then
 * reports what the context's device status word recorded since the last
 * synchronizing call: LAIR_ERR_SYNC when a cooperative kernel's bounded wait
 * gave up, LAIR_ERR_SHAPE when a caller-supplied pivot was outside [0, n).
 * The word is sticky until a synchronizing call (this one, or any call
 * documented as synchronizing) reads and resets it; while it is set, the
 * cooperative kernels give up their waits at once.  No reference analogue. */
lair_status lair_context_synchronize(lair_context *ctx);
/* Polls a cooperative kernel's bounded wait makes before it gives up
 * (default 8388608, about seconds).  0 makes every wait give up at once, so
 * that the calls that use such kernels (getrf of more than one column
 * panel, getrs and the adjoint for n > 128, the mixed solver) return
 * LAIR_ERR_SYNC deterministically: used to test the failure path.  No
 * reference analogue. */
lair_status lair_context_set_wait_limit(lair_context *ctx, int64_t polls);
/* The HIP stream (hipStream_t) the context launches on, as an opaque pointer. */
void *lair_context_stream(lair_context *ctx);
/* Human-readable message for the last error seen on this thread. */
const char *lair_last_error(void);

/* Device memory helpers, so callers need no other GPU runtime.  The copies
 * block until the data has arrived. */
lair_status lair_malloc(void **dptr, uint64_t bytes);
lair_status lair_free(void *dptr);
lair_status lair_memcpy_h2d(void *dst, const void *src, uint64_t bytes);
lair_status lair_memcpy_d2h(void *dst, const void *src, uint64_t bytes);
lair_status lair_memcpy_d2d(void *dst, const void *src, uint64_t bytes);

/* ------------------------------------------- LAPACK-level entry points --- */

/* Replaces lapack::getrf (src/lapack/getrf.rs:12-27): in-place LU with
 * partial pivoting of the m x n matrix (a, rs, cs); ipiv receives min(m,n)
 * pivots and *singular the last zero-pivot column or -1.  Pivots match the
 * reference's row-major kernel (getrf.rs:46-120) exactly, apart from choices
 * between values that differ only by rounding. */
lair_status lair_dgetrf(int64_t m, int64_t n, double *a, int64_t rs, int64_t cs, int64_t *ipiv,
                        int64_t *singular);
lair_status lair_sgetrf(int64_t m, int64_t n, float *a, int64_t rs, int64_t cs, int64_t *ipiv,
                        int64_t *singular);

/* Replaces lapack::getrs (src/lapack/getrs.rs:12-39), extended to nrhs
 * right-hand sides: solves A X = B (trans = 0) or A^T X = B (trans = 1) from
 * the packed factors (lu, lrs, lcs) and pivots; b (n x nrhs, strides brs, bcs)
 * is overwritten with X.  Like the reference, a shape mismatch is the caller's
 * error; it returns LAIR_ERR_SHAPE here instead of panicking. */
lair_status lair_dgetrs(int64_t n, int64_t nrhs, const double *lu, int64_t lrs, int64_t lcs,
                        const int64_t *ipiv, double *b, int64_t brs, int64_t bcs, int32_t trans);
lair_status lair_sgetrs(int64_t n, int64_t nrhs, const float *lu, int64_t lrs, int64_t lcs,
                        const int64_t *ipiv, float *b, int64_t brs, int64_t bcs, int32_t trans);

/* Replaces lapack::laswp (src/lapack/laswp.rs:11-40): rows i <-> ipiv[i] for
 * i in begin..npiv, in forward order, over ncols columns. */
lair_status lair_dlaswp(int64_t ncols, double *a, int64_t rs, int64_t cs, int64_t begin,
                        int64_t npiv, const int64_t *ipiv);

/* Replaces equation::solve (src/equation.rs:32-60): x = A^-1 b for square A
 * (n x n, strides rs, cs) and b of length blen.  Returns LAIR_ERR_SHAPE when A
 * is not square (m != n) or blen != m, LAIR_ERR_SINGULAR when the
 * factorization met an exactly zero pivot.  a and b are not modified. */
lair_status lair_dsolve(int64_t m, int64_t n, const double *a, int64_t rs, int64_t cs,
                        int64_t blen, const double *b, double *x);

/* Complex<f64> (z) and Complex<f32> (c) forms of getrf, getrs and solve: the
 * reference's generic A: Scalar instances for Complex<T> (src/scalar.rs:
 * 310-330), with interleaved (re, im) storage as num-complex 0.4 lays out
 * Complex<T>; strides count complex elements.  Pivoting picks the first
 * maximum of |re| + |im| (src/blas/iamax.rs:14); the pivot reciprocal, the
 * products and getrs's division are num-complex's (getrf.rs:76, getrs.rs:34).
 * getrs has no transposed form, like the reference's.  Limits: m <= 32768
 * rows for zgetrf (65536 for cgetrf), n <= 16384 for getrs; beyond them
 * LAIR_ERR_UNSUPPORTED. */
lair_status lair_zgetrf(int64_t m, int64_t n, double *a, int64_t rs, int64_t cs, int64_t *ipiv,
                        int64_t *singular);
lair_status lair_cgetrf(int64_t m, int64_t n, float *a, int64_t rs, int64_t cs, int64_t *ipiv,
                        int64_t *singular);
lair_status lair_zgetrs(int64_t n, int64_t nrhs, const double *lu, int64_t lrs, int64_t lcs,
                        const int64_t *ipiv, double *b, int64_t brs, int64_t bcs);
lair_status lair_cgetrs(int64_t n, int64_t nrhs, const float *lu, int64_t lrs, int64_t lcs,
                        const int64_t *ipiv, float *b, int64_t brs, int64_t bcs);
lair_status lair_zsolve(int64_t m, int64_t n, const double *a, int64_t rs, int64_t cs,
                        int64_t blen, const double *b, double *x);
lair_status lair_csolve(int64_t m, int64_t n, const float *a, int64_t rs, int64_t cs,
                        int64_t blen, const float *b, float *x);

/* ----------------------------------------------- device-resident calls --- */

/* getrf on a device-resident row-major matrix with leading dimension lda.
 * nb = panel width: > 0 that width (capped at 256), 0 the default (128, or
 * 64 when min(m,n) <= 256), < 0 autotuned: widths 64, 128 and 256 each factor
 * a scratch copy once and the fastest is kept per shape in the context (the
 * first such call costs three extra factorizations).  Widths change only the
 * summation order of the trailing updates.  ipiv is a device int64 array of min(m,n)
 * entries.  *singular (host) is written after the call completes (the call
 * synchronizes the stream when singular != NULL). */
lair_status lair_dgetrf_device(lair_context *ctx, int64_t m, int64_t n, double *a, int64_t lda,
                               int64_t *ipiv, int64_t nb, int64_t *singular);
lair_status lair_sgetrf_device(lair_context *ctx, int64_t m, int64_t n, float *a, int64_t lda,
                               int64_t *ipiv, int64_t nb, int64_t *singular);
/* FP32 LU on device data with a choice of pivoting and trailing-update
 * arithmetic, as used inside the mixed solver: `pivoting` is 0 (partial
 * pivoting, as lair_sgetrf_device) or LAIR_PIVOT_TOURNAMENT; `prec` is
 * LAIR_PREC_FP32, LAIR_PREC_BF16X6 or LAIR_PREC_BF16.  ipiv follows the
 * LAPACK convention in both cases, so P A = L U with P from ipiv.  No
 * reference counterpart (the reference has partial pivoting only). */
lair_status lair_sgetrf_pivot_device(lair_context *ctx, int64_t m, int64_t n, float *a, int64_t lda,
                                     int64_t *ipiv, int64_t nb, int32_t pivoting, int32_t prec,
                                     int64_t *singular);

/* getrs on device data: b is n x nrhs row-major (leading dimension ldb),
 * overwritten with the solution of A X = B (trans = 0) or A^T X = B.  The
 * pivots are range-checked on the device; the call synchronizes and returns
 * LAIR_ERR_SHAPE when one is outside [0, n), as the host-pointer getrs does. */
lair_status lair_dgetrs_device(lair_context *ctx, int64_t n, int64_t nrhs, const double *lu,
                               int64_t lda, const int64_t *ipiv, double *b, int64_t ldb,
                               int32_t trans);
lair_status lair_sgetrs_device(lair_context *ctx, int64_t n, int64_t nrhs, const float *lu,
                               int64_t lda, const int64_t *ipiv, float *b, int64_t ldb,
                               int32_t trans);

/* C -= A B on device data: the reference's gemm (src/blas/gemm.rs:6-32) with
 * alpha = -1 and no conjugation, on the trailing-update kernels of getrf.  A
 * is m x k (leading dimension lda >= k), B is k x n (ldb >= n), C is m x n
 * (ldc >= n), all row-major.  The single-precision form takes prec:
 * LAIR_PREC_FP32 (FP32 MFMA), LAIR_PREC_BF16X6 (six BF16 products, FP32
 * accumulation) or LAIR_PREC_BF16 (one BF16 product).  The sums run in an
 * order of the tiles' choosing, so results match the reference's loop to
 * rounding, not bit for bit.  Returns LAIR_ERR_UNSUPPORTED for a leading
 * dimension beyond 2^31 bytes per 128 rows.  Synchronizes the stream. */
lair_status lair_dgemm_minus_device(lair_context *ctx, int64_t m, int64_t n, int64_t k, const double *a,
                                    int64_t lda, const double *b, int64_t ldb, double *c, int64_t ldc);
lair_status lair_sgemm_minus_device(lair_context *ctx, int32_t prec, int64_t m, int64_t n, int64_t k,
                                    const float *a, int64_t lda, const float *b, int64_t ldb, float *c,
                                    int64_t ldc);

/* Complex device forms (interleaved storage, lda/ldb in complex elements,
 * pointers aligned to the complex element).  nb: panel width, 0 or < 0 the
 * default (64, the widest complex panel). */
lair_status lair_zgetrf_device(lair_context *ctx, int64_t m, int64_t n, double *a, int64_t lda,
                               int64_t *ipiv, int64_t nb, int64_t *singular);
lair_status lair_cgetrf_device(lair_context *ctx, int64_t m, int64_t n, float *a, int64_t lda,
                               int64_t *ipiv, int64_t nb, int64_t *singular);
lair_status lair_zgetrs_device(lair_context *ctx, int64_t n, int64_t nrhs, const double *lu,
                               int64_t lda, const int64_t *ipiv, double *b, int64_t ldb);
lair_status lair_cgetrs_device(lair_context *ctx, int64_t n, int64_t nrhs, const float *lu,
                               int64_t lda, const int64_t *ipiv, float *b, int64_t ldb);

/* lair_dgetrs_device without any device allocation: no counterpart in the
 * reference.  For callers that cannot allocate (code under stream capture,
 * tight serving loops).  The workspace (row permutation, inverted diagonal
 * blocks, staging copy, tagged solved values) lives in a static device array
 * of 22.1 MiB held by every GPU the library runs on for the life of the
 * process; n <= 16384 and nrhs <= 16, larger calls return
 * LAIR_ERR_UNSUPPORTED.  Every call on one GPU shares that array, so calls
 * that overlap on one GPU (two contexts, two host threads) race; calls on
 * one context are serialized.  Same results as lair_dgetrs_device;
 * synchronizes.  With profiling enabled (lair_profile_enable) the timing
 * events it records are created on first use. */
lair_status lair_dgetrs_device_noalloc(lair_context *ctx, int64_t n, int64_t nrhs, const double *lu,
                                       int64_t lda, const int64_t *ipiv, double *b, int64_t ldb,
                                       int32_t trans);

/* Result record of the mixed-precision solver. */
typedef struct lair_ir_info {
  int32_t iterations;   /* refinement steps taken */
  int32_t converged;    /* 1 when the backward error met the tolerance */
  int32_t fell_back;    /* 1 when the FP64 factorization was used instead */
  int32_t precision;    /* LAIR_PREC_* actually used for the final answer */
  double backward_error;   /* max over columns of ||b-Ax||inf / (||A||inf ||x||inf) */
  double history[32];      /* backward error after each step (0: initial solve) */
} lair_ir_info;

/* Mixed-precision solve A X = B from device-resident FP64 data (row-major,
 * lda, ldb, ldx): the factorization runs in `prec` (LAIR_PREC_FP32 /
 * LAIR_PREC_BF16X6 / LAIR_PREC_BF16, optionally | LAIR_PIVOT_TOURNAMENT;
 * LAIR_PREC_FP64 runs the FP64 solver) and FP64 accuracy is recovered by
 * iterative refinement with the residual, correction update and norms in FP64.
 * Stops when the normwise backward error is <= tol (tol <= 0 selects
 * n * DBL_EPSILON) or after max_iter steps; a system that does not converge is
 * re-solved with the FP64 factorization (info->fell_back = 1, status
 * LAIR_ERR_NOCONV).  A and B are not modified; X must not overlap them.
 * Returns when all of its work on the context's stream has finished (the
 * refinement reads its norms from pinned host memory that a kernel writes
 * last, or synchronizes the stream). */
lair_status lair_dgesv_mixed_device(lair_context *ctx, int64_t n, int64_t nrhs, const double *a,
                                    int64_t lda, const double *b, int64_t ldb, double *x,
                                    int64_t ldx, int32_t prec, int32_t max_iter, double tol,
                                    lair_ir_info *info);

/* FP64 solve A X = B from device data (getrf + getrs), A and B untouched.
 * The factors are kept in the context's workspace.  Synchronizes the stream
 * and returns LAIR_ERR_SINGULAR for an exactly singular matrix. */
lair_status lair_dgesv_device(lair_context *ctx, int64_t n, int64_t nrhs, const double *a,
                              int64_t lda, const double *b, int64_t ldb, double *x, int64_t ldx);

/* Normwise backward error of X for A X = B, all device FP64 data:
 * max over columns of ||b - A x||inf / (||A||inf ||x||inf).  Synchronizes. */
lair_status lair_dbackward_error(lair_context *ctx, int64_t n, int64_t nrhs, const double *a,
                                 int64_t lda, const double *b, int64_t ldb, const double *x,
                                 int64_t ldx, double *berr);

/* ------------------------------------------------------------ batched --- */

/* Batched solve of `count` independent systems A_i X_i = B_i with
 * 1 <= n_i <= 256 and 1 <= nrhs_i <= 16 (equation::solve, src/equation.rs:32-60,
 * applied to a batch; no reference counterpart for the batching itself).
 * Packed device layout, systems in order: A_i row-major n_i x n_i at element
 * offset sum_{j<i} n_j^2 of `a`; B_i row-major n_i x nrhs_i at offset
 * sum_{j<i} n_j nrhs_j of `b`; pivots of A_i (0-based, int64) at offset
 * sum_{j<i} n_j of `ipiv`; info[i] = -1 or the last zero-pivot column.
 * a is overwritten by the LU factors and b by X.  precision: 0 = FP64,
 * 1 = FP32 (a, b are float).  The plan sorts the systems by decreasing n,
 * optionally autotunes (flags & LAIR_BATCHED_AUTOTUNE: the systems may be
 * split into two size classes, and each class's launch shape is chosen by
 * timing candidates on scratch copies of a and b) and captures the launches
 * in one hipGraph on the context's stream. */
#define LAIR_BATCHED_AUTOTUNE 1
typedef struct lair_batched_plan lair_batched_plan;
lair_status lair_gesv_batched_plan(lair_context *ctx, int32_t precision, int32_t count,
                                   const int32_t *n, const int32_t *nrhs, void *a, void *b,
                                   int64_t *ipiv, int32_t *info, int32_t flags,
                                   lair_batched_plan **plan);
/* Replays the captured graph (asynchronous on the context's stream). */
lair_status lair_gesv_batched_run(lair_batched_plan *plan);
/* Launch shape chosen for the largest systems: panel width, workgroups, LDS
 * budget (elements). */
lair_status lair_gesv_batched_config(lair_batched_plan *plan, int32_t *nb, int32_t *grid,
                                     int32_t *lds_elems);
lair_status lair_gesv_batched_destroy(lair_batched_plan *plan);

/* ------------------------------------------------------------ adjoint --- */

/* Adjoint of X = A^-1 B (no counterpart in the reference; the reverse-mode
 * derivative of equation::solve, src/equation.rs:32-60).  From the factors
 * (lu, ipiv) of A, the solution X (n x nrhs) and the upstream gradient XBAR:
 *   BBAR = A^-T XBAR   (transposed getrs on the stored factors)
 *   ABAR = -BBAR X^T   (n x n, FP64 MFMA GEMM, written without reading ABAR)
 * All arrays are device-resident and row-major.  Synchronizes the stream. */
lair_status lair_dsolve_backward_device(lair_context *ctx, int64_t n, int64_t nrhs,
                                        const double *lu, int64_t lda, const int64_t *ipiv,
                                        const double *x, int64_t ldx, const double *xbar,
                                        int64_t ldxb, double *bbar, int64_t ldbb, double *abar,
                                        int64_t ldab);
/* Host-pointer form with contiguous row-major arrays (lu n x n, x/xbar/bbar
 * n x nrhs, abar n x n). */
lair_status lair_dsolve_backward(int64_t n, int64_t nrhs, const double *lu, const int64_t *ipiv,
                                 const double *x, const double *xbar, double *bbar, double *abar);

/* ------------------------------------------------ device-resident LU --- */

/* lu::Factorized (src/decomposition/lu.rs:12-20) with its factors kept on the
 * device: the packed LU, the pivots, the row permutation the factorization
 * built and, for square systems (from the first solve on), the inverses of
 * the diagonal blocks of L and U, so that repeated solves and adjoints reuse
 * them.  FP64. */
typedef struct lair_lu lair_lu;

/* Replaces From<ArrayBase> for Factorized (lu.rs:156-171): factors a copy of
 * the device matrix a (m x n, row-major, lda); nb as lair_dgetrf_device.
 * Synchronizes. */
lair_status lair_dlu_factor(lair_context *ctx, int64_t m, int64_t n, const double *a, int64_t lda, int64_t nb,
                            lair_lu **lu);
lair_status lair_dlu_destroy(lair_lu *lu);
/* Shape and singular column (-1 for None): Factorized::is_singular
 * (lu.rs:75-77) is singular >= 0. */
lair_status lair_dlu_info(const lair_lu *lu, int64_t *m, int64_t *n, int64_t *singular);
/* Device pointers of the packed factors (leading dimension ld), the pivots
 * (min(m, n) entries, LAPACK convention) and the row permutation
 * (perm[i] = original row now at row i).  Owned by the handle. */
lair_status lair_dlu_factors(const lair_lu *lu, const double **lu_dev, int64_t *ld, const int64_t **ipiv,
                             const int32_t **perm);
/* Replaces Factorized::solve (lu.rs:87-98), extended to nrhs columns and to
 * A^T X = B (trans = 1): b (device, n x nrhs, ldb) is overwritten with X.
 * LAIR_ERR_SHAPE for a non-square factorization.  Synchronizes. */
lair_status lair_dlu_solve(lair_context *ctx, lair_lu *lu, int64_t nrhs, double *b, int64_t ldb, int32_t trans);
/* lair_dsolve_backward_device on the handle's factors (no permutation
 * rebuild, no diagonal re-inversion).  Synchronizes. */
lair_status lair_dlu_solve_backward(lair_context *ctx, lair_lu *lu, int64_t nrhs, const double *x, int64_t ldx,
                                    const double *xbar, int64_t ldxb, double *bbar, int64_t ldbb,
                                    double *abar, int64_t ldab);
/* Factor extraction on the device into out (row-major, ldo):
 * LAIR_LU_P  m x m: P with P[perm[j]][j] = 1          (Factorized::p, lu.rs:28-39)
 * LAIR_LU_L  m x k: unit lower, k = min(m, n)          (Factorized::l, lu.rs:42-57)
 * LAIR_LU_U  k x n: upper                              (Factorized::u, lu.rs:60-72)
 * LAIR_LU_PL m x k: P * L                              (Factorized::into_pl, lu.rs:107-153)
 * Synchronizes. */
#define LAIR_LU_P 0
#define LAIR_LU_L 1
#define LAIR_LU_U 2
#define LAIR_LU_PL 3
lair_status lair_dlu_extract(lair_context *ctx, const lair_lu *lu, int32_t which, double *out, int64_t ldo);

/* ----------------------------------------------- instrumentation ------- */

/* Kernel classes timed with HIP events on the context's stream while
 * profiling is enabled. */
#define LAIR_KCLASS_PANEL 0
#define LAIR_KCLASS_LASWP 1
#define LAIR_KCLASS_TRSM 2
#define LAIR_KCLASS_GEMM 3 /* trailing-matrix update A22 -= L21 U12 */
#define LAIR_KCLASS_GETRS 4
#define LAIR_KCLASS_RESIDUAL 5
#define LAIR_KCLASS_CONVERT 6
#define LAIR_KCLASS_GEMM_PANEL 7 /* updates inside a panel (narrow K) */
#define LAIR_KCLASS_COUNT 8

/* class_mask: bit k enables class k; 0 disables; LAIR_PROFILE_ALL enables all. */
#define LAIR_PROFILE_ALL 0xFF
lair_status lair_profile_enable(lair_context *ctx, int32_t class_mask);
/* Accumulated milliseconds, launch count, algorithmic flops and algorithmic
 * bytes (minimum HBM traffic of the launches' operands) for a class since the
 * last reset (synchronizes the stream). */
lair_status lair_profile_read(lair_context *ctx, int32_t kclass, double *ms, int64_t *launches,
                              double *flops, double *bytes);
lair_status lair_profile_reset(lair_context *ctx);

#ifdef __cplusplus
}
#endif

#endif /* LAIR_DEV_H */
