/*
 * mat_qr.h -- C ABI of the MI355X (gfx950) Householder QR library.
 *
 * The reference (Mat-8.3.1) is a header-only C++ library with no FFI layer:
 * its "operator API" is the template call QR(A, Q, R, QR_Householder)
 * (QR/QR.h:7-46) on column-major MAT<T> (Base/__Matrix.h:2-174).  This header
 * is the boundary that call is routed through: plain pointers, 64-bit sizes,
 * explicit leading dimensions, no C++ or HIP types.  include/mat/Mat.h keeps
 * the reference's MAT<T> / QR() signatures on top of it.
 *
 * Conventions
 *  - Column-major storage, element (r, c) at A[r + c * lda], lda >= max(1, m).
 *  - Every function returns an int status: 0 = success; -i = the i-th
 *    argument is invalid (LAPACK info style, counting from 1); positive codes
 *    are runtime errors (MQR_ERR_*), described by mqr_strerror().  No C++
 *    exception crosses this boundary.
 *  - Functions taking a handle run on the handle's device and HIP stream and
 *    are stream-ordered: they return once the work is queued unless the
 *    description says otherwise.  Device pointers must be device memory of the
 *    handle's device.  Distinct handles may be used from distinct threads.
 *  - Reflector conventions follow the reference panel
 *    (QR/__Householder_Blocked.h:62-84): R_jj = -sign(a_jj) * ||a_j:m,j||
 *    with sign(0) = +, v_j = 1 implicit, tau_j = (|a_jj| + ||.||) / ||.||;
 *    a column whose norm is 0 gets tau = 0; a column with nothing below the
 *    diagonal is still reflected (tau = 2).
 *  - Compact-WY T blocks (forward, columnwise) use the LAPACK dgeqrt layout:
 *    for the block of nb columns starting at column j, T(0:b, j:j+b) is the
 *    b x b upper-triangular factor with I - V T V^T = H_j ... H_{j+b-1}.
 *    The reference keeps the same factors at T_all + j * BS with leading
 *    dimension b (QR/__Householder_Blocked.h:236).
 */
#ifndef MAT_QR_H
#define MAT_QR_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------ */
#define MQR_SUCCESS 0
#define MQR_ERR_NO_DEVICE 1001     /* no HIP device / device index out of range */
#define MQR_ERR_HIP 1002           /* a HIP runtime call failed */
#define MQR_ERR_ALLOC 1003         /* device allocation failed */
#define MQR_ERR_NOT_CONVERGED 1004 /* iterative refinement did not converge */
#define MQR_ERR_UNSUPPORTED 1005   /* valid request this build does not support */
#define MQR_ERR_NOT_INITIALIZED 1006
#define MQR_ERR_INTERNAL 1007
#define MQR_ERR_TOO_LARGE 1008     /* request exceeds a fixed-size device scratch array */
#define MQR_ERR_IO 1009            /* a file could not be opened / read / written, or its
                                      size or header disagrees with the request */

typedef struct mqr_handle_st *mqr_handle_t;

typedef enum { MQR_F64 = 0, MQR_F32 = 1 } mqr_dtype_t;

/* Arithmetic of the trailing updates in the mixed-precision least squares.
 * Panels, T factors and refinement residuals always stay FP64. */
typedef enum {
    MQR_PREC_FP64 = 0,   /* plain FP64 MFMA */
    MQR_PREC_FP32 = 1,   /* FP32 MFMA, FP32 accumulation */
    MQR_PREC_BF16X3 = 2, /* split-BF16 (hi*hi + hi*lo + lo*hi) MFMA, FP32 accumulation:
                            ~2^-16 per product, converges for kappa up to ~1e4 */
    MQR_PREC_BF16 = 3    /* single BF16 MFMA product, FP32 accumulation: ~2^-8,
                            converges for kappa up to ~1e2; beyond, refinement
                            stagnates and the FP64 fallback runs */
} mqr_precision_t;

/* ---- library / handle ---------------------------------------------------- */
/* No reference counterpart (the reference has no runtime state). */
const char *mqr_strerror(int status);
int mqr_get_device_count(int *count);
int mqr_create(mqr_handle_t *handle, int device);
int mqr_destroy(mqr_handle_t handle);
/* stream: a hipStream_t; NULL selects the default (null) stream.  A new
 * handle starts on a non-blocking stream of its own (mqr_get_stream). */
int mqr_set_stream(mqr_handle_t handle, void *stream);
int mqr_get_stream(mqr_handle_t handle, void **stream);
/* Block size nb of the blocked factorisation (the reference's __BlockSize_QR,
 * Mat.h:79-81,105-107): 16, 32, 48 or 64 (panels of nb columns), or 128,
 * 192 or 256 (two-level blocking: panels of 64 columns, each nb-column block
 * applied to the far columns as one block reflector; the T blocks are then
 * nb x nb).  The default 64 also applies groups of four panels to the far
 * columns at once, keeping its T blocks 64 wide. */
int mqr_set_block_size(mqr_handle_t handle, int nb);
int mqr_get_block_size(mqr_handle_t handle, int *nb);
int mqr_synchronize(mqr_handle_t handle);
/* Device memory helpers so that callers need no HIP headers. */
int mqr_malloc(mqr_handle_t handle, void **ptr, size_t bytes);
int mqr_free(mqr_handle_t handle, void *ptr);
int mqr_memcpy_htod(mqr_handle_t handle, void *dst, const void *src, size_t bytes);
int mqr_memcpy_dtoh(mqr_handle_t handle, void *dst, const void *src, size_t bytes);

/* ---- kernel timing (no reference counterpart; the reference has no profiler) */
typedef struct {
    char name[24];    /* kernel kind: panel_gemm, vta_gemm, panel_small, w2, update_gemm, ... */
    int64_t launches; /* launches timed */
    double total_ms;  /* sum of HIP-event durations */
    double flops;     /* algorithmic flops of those launches */
    double bytes;     /* algorithmic HBM bytes of those launches */
} mqr_kernel_stat_t;

/* Kernel kinds, in the order of their bit in the mask below. */
typedef enum {
    MQR_KIND_PANEL_GEMM = 0,  /* GEMMs and reductions inside the panel */
    MQR_KIND_VTA_GEMM = 1,    /* split-K W = V^T C */
    MQR_KIND_PANEL_SMALL = 2, /* single-workgroup panel kernels */
    MQR_KIND_W2 = 3,          /* Y = V T^T and other T products */
    MQR_KIND_UPDATE_GEMM = 4, /* trailing update C -= Y W */
    MQR_KIND_TRSM = 5,        /* triangular solves and inverses */
    MQR_KIND_GEMM = 6,        /* mqr_dgemm / mqr_sgemm */
    MQR_KIND_RESID = 7,       /* refinement residual passes */
    MQR_KIND_CAST = 8,        /* precision conversions */
    MQR_KIND_OTHER = 9,
    MQR_KIND_LOOKAHEAD = 10,  /* next-block update on the critical stream */
    MQR_KIND_APPLY_Q = 11,    /* Q / Q^T applications (ormqr, orgqr) */
    MQR_KIND_NEAR_UPDATE = 12 /* a panel's update of the columns up to the next outer block's end */
} mqr_kernel_kind_t;

/* Turn per-launch HIP event timing on or off for calls on this handle
 * (clears what was recorded). */
int mqr_profile_enable(mqr_handle_t handle, int on);
/* Time only the kinds whose bit (1u << kind) is set (default: all).  Two
 * events per timed launch add to the stream's dispatch gaps; timing one kind
 * keeps that overhead small. */
int mqr_profile_set_mask(mqr_handle_t handle, unsigned mask);
/* Synchronise, then return one entry per kernel kind timed since the last
 * read or enable (at most max_stats are written; *count gets the number of
 * kinds), and clear the records. */
int mqr_profile_read(mqr_handle_t handle, mqr_kernel_stat_t *stats, int max_stats, int *count);

/* ---- FP64 GEMM tile (device pointers) ----------------------------------- */
/* C = alpha * op(A) * op(B) + beta * C with op(X) = X ('N') or X^T ('T');
 * op(A) is m x k, op(B) k x n.  The device counterpart of the packed
 * MR x NR micro-kernel and its drivers behind operator*
 * (Mul/__micro_kernel_packed_double.h:301-585, Mul/Mul.h:123-235). */
int mqr_dgemm(mqr_handle_t handle, char transa, char transb, int64_t m, int64_t n, int64_t k,
              double alpha, const double *A, int64_t lda, const double *B, int64_t ldb,
              double beta, double *C, int64_t ldc);

/* Same for float (FP32 MFMA, FP32 accumulation): the counterpart of the
 * generic packed micro-kernel for float (Mul/__micro_kernel_packed.h:1-307). */
int mqr_sgemm(mqr_handle_t handle, char transa, char transb, int64_t m, int64_t n, int64_t k,
              float alpha, const float *A, int64_t lda, const float *B, int64_t ldb, float beta,
              float *C, int64_t ldc);

/* The same product on the BF16 MFMA with FP32 accumulation: every FP32
 * operand x is split as hi = bf16(x), lo = bf16(x - hi); terms = 1 multiplies
 * hi * hi, terms = 3 adds hi * lo + lo * hi (relative error ~2^-16).  The
 * precision the mixed solve's MQR_PREC_BF16 / MQR_PREC_BF16X3 updates use.
 * No reference counterpart (the reference has no BF16 path). */
int mqr_sgemm_bf16(mqr_handle_t handle, int terms, char transa, char transb, int64_t m, int64_t n,
                   int64_t k, float alpha, const float *A, int64_t lda, const float *B,
                   int64_t ldb, float beta, float *C, int64_t ldc);

/* Host-memory products behind MAT<T>::operator* (Mul/Mul.h:205-235):
 * C (m x n) = A (m x k) * B (k x n), all column-major with leading dimensions
 * m, k, m; the operands are copied to the device, multiplied on the MFMA
 * tiles above and C is copied back.  Blocks until C is written. */
int mqr_dgemm_host(mqr_handle_t handle, int64_t m, int64_t n, int64_t k, const double *A,
                   const double *B, double *C);
int mqr_sgemm_host(mqr_handle_t handle, int64_t m, int64_t n, int64_t k, const float *A,
                   const float *B, float *C);

/* ---- FP64 blocked Householder QR (device pointers) ------------------------ */
/* Factor the m x n matrix A in place: R on and above the diagonal, the
 * reflectors below it, tau[min(m,n)], and the T blocks in T (ldt >= nb, ldt x
 * min(m,n)).  Replaces __QR_HouseHolder_Blocked(R, buf, m, n)
 * (QR/__Householder_Blocked.h:214-245) with its panel (:62-84), T build
 * (:86-103) and trailing update (:143-212). */
int mqr_dgeqrf(mqr_handle_t handle, int64_t m, int64_t n, double *A, int64_t lda,
               double *tau, double *T, int64_t ldt);

/* Q(:, 0:nq) = H_0 ... H_{k-1} applied to the first nq columns of I_m
 * (nq = k: thin Q, nq = m: full Q), from a factorisation by mqr_dgeqrf.
 * Replaces __QR_HouseHolder_Blocked_CalculateQ (QR/__Householder_Blocked.h:247-333);
 * the known-zero part of Q is skipped.
 * T holds one block per nb columns, with the handle's block size nb at the
 * time of mqr_dgeqrf: keep nb unchanged between the factorisation and the Q
 * formation or application.  A T that one of the handle's last 16
 * factorisations wrote with another nb is rejected (-7 here, -8 for
 * mqr_dormqr). */
int mqr_dorgqr(mqr_handle_t handle, int64_t m, int64_t nq, int64_t k, const double *VR,
               int64_t ldvr, const double *T, int64_t ldt, double *Q, int64_t ldq);

/* C <- Q^T C (trans = 'T') or Q C (trans = 'N') for the m x nrhs matrix C,
 * with Q held as k reflectors by mqr_dgeqrf.  The block-reflector pattern of
 * QR/__Householder_Blocked.h:143-212 with C as the right-hand side; T as
 * for mqr_dorgqr (same block size as the factorisation). */
int mqr_dormqr(mqr_handle_t handle, char trans, int64_t m, int64_t nrhs, int64_t k,
               const double *VR, int64_t ldvr, const double *T, int64_t ldt, double *C,
               int64_t ldc);

/* Solve R X = B (trans = 'N') or R^T X = B (trans = 'T') in place for the
 * upper triangle of the n x n matrix R (the diagonal must be nonzero).
 * Pattern of the blocked back substitution LU/__Solve_direct.h:39-65. */
int mqr_dtrsm_upper(mqr_handle_t handle, char trans, int64_t n, int64_t nrhs, const double *R,
                    int64_t ldr, double *B, int64_t ldb);

/* Plain FP64 least squares min ||A X - B|| for m >= n: A is overwritten by
 * its factorisation, B (m x nrhs) by Q^T B, whose first n rows hold X.
 * No reference counterpart (new on top of the QR path). */
int mqr_dgels(mqr_handle_t handle, int64_t m, int64_t n, int64_t nrhs, double *A, int64_t lda,
              double *B, int64_t ldb);

/* ---- the other QR_type values of QR(A, Q, R, type) (QR/QR.h:26-42) ------- */
/* Givens rotations (QR/__Givens.h:45-63): A m x n (lda) is read, Q (m x m,
 * ldq) and R (m x n, ldr) are written; R may be A (in place).  The
 * reference's rotation order and arithmetic, so Q and R equal the
 * reference's bit for bit; R keeps the rotations' rounding residue below the
 * diagonal like the reference.  Where the reference divides by a zero pivot
 * (filling R and Q with NaN) the rotation takes its limit instead (c = 0,
 * s = sign of the entry below; the identity when both are zero).  One step of
 * the sweep holds at most 8192 rotations: min(n, m / 2 + 1) <= 8192, else
 * MQR_ERR_UNSUPPORTED. */
int mqr_dgeqr_givens(mqr_handle_t handle, int64_t m, int64_t n, const double *A, int64_t lda,
                     double *Q, int64_t ldq, double *R, int64_t ldr);
/* Gram-Schmidt (QR/__GramSchmidt.h:2-18), m >= n: the thin QR with a
 * positive diagonal, A = Q R with Q m x n (ldq) orthonormal and R n x n
 * (ldr >= n) upper triangular, zero below the diagonal.  Q may be A (in
 * place).  Computed by block Gram-Schmidt with reorthogonalisation, so Q is
 * orthogonal to working precision (the reference's modified Gram-Schmidt
 * loses orthogonality in proportion to cond(A)); both give the same unique
 * factors up to that rounding. */
int mqr_dgeqr_gramschmidt(mqr_handle_t handle, int64_t m, int64_t n, const double *A, int64_t lda,
                          double *Q, int64_t ldq, double *R, int64_t ldr);

/* ---- LU and the direct solve (LU/__LU.h, LU/__PLU.h, LU/__Solve_direct.h) */
/* Blocked LU of A (m x n, lda) in place: L unit lower (strictly below the
 * diagonal), U on and above.  ipiv (device int32[min(m, n)]) selects partial
 * pivoting, the first row of largest magnitude as in PLU (LU/__PLU.h:18-50):
 * row i was exchanged with row ipiv[i] (0-based, in order); NULL factors
 * without pivoting as LU (LU/__LU.h:2-16).  info (device int32, may be
 * NULL): 0, or the 1-based column of the first exactly zero pivot, where the
 * reference stops (`break`); here the factorisation goes on LAPACK-style. */
int mqr_dgetrf(mqr_handle_t handle, int64_t m, int64_t n, double *A, int64_t lda, int32_t *ipiv,
               int32_t *info);
/* B (n x nrhs, ldb) <- A^{-1} B from mqr_dgetrf's factors of the n x n A
 * (ipiv NULL when it factored without pivoting). */
int mqr_dgetrs(mqr_handle_t handle, int64_t n, int64_t nrhs, const double *LU, int64_t ldlu,
               const int32_t *ipiv, double *B, int64_t ldb);
/* Solve_direct (LU/__Solve_direct.h:175-220): A (n x n) is overwritten by
 * its PLU factors, ipiv (device, n) by the pivots, B (n x nrhs) by X. */
int mqr_dgesv(mqr_handle_t handle, int64_t n, int64_t nrhs, double *A, int64_t lda, int32_t *ipiv,
              double *B, int64_t ldb, int32_t *info);

/* ---- mixed precision least squares with iterative refinement ------------ */
typedef struct {
    int max_iter;    /* refinement steps allowed (default 30) */
    double tol;      /* stop when the backward-error estimate <= tol (default 1e-14);
                        the preconditioned CGLS stage stops at tol / 16 */
    int fallback;    /* 1: on stagnation refactor in plain FP64 and solve (default 1) */
    int krylov_max_iter; /* preconditioned CGLS steps allowed once the stationary
                            refinement stalls (0: default 400; < 0: none) */
} mqr_ir_opts_t;

typedef struct {
    int iterations;       /* refinement steps taken */
    int converged;        /* 1 when tol was met */
    int fell_back;        /* 1 when the FP64 fallback ran */
    double backward_error;/* ||R^{-T} A^T r|| / (||A||_F ||x||) at exit: bounds the
                             Karlson-Walden normwise backward error from above */
    double residual_norm; /* ||b - A x|| at exit */
    int krylov_iterations;/* preconditioned CGLS steps taken (included in iterations) */
} mqr_ir_info_t;

/* x = argmin ||A x - b|| for one right-hand side, m >= n.  A (m x n, FP64,
 * device) and b (m, device) are not modified; x (n, device) is written.
 * The factorisation runs its trailing updates in `prec` (panels, T factors
 * and R stay FP64).  Refinement with FP64 residuals r = b - A x from the
 * caller's A: first corrected semi-normal equations (x += R^{-1} R^{-T} A^T r,
 * converging while u_prec * cond(A) < 1); if that stalls, conjugate
 * gradients on the least-squares problem preconditioned by R (CGLS on
 * A R^{-1}, whose condition number is about 1 + u_prec * cond(A)), which
 * carries split-BF16 factors to FP64 accuracy at cond(A) = 1e6; if that
 * stalls too, the plain FP64 solve runs instead when opts->fallback is set
 * (info->fell_back).  opts may be NULL (defaults).
 * Blocks until the solve is done (the stopping test is read on the host).
 * No reference counterpart. */
int mqr_gels_mixed(mqr_handle_t handle, mqr_precision_t prec, int64_t m, int64_t n,
                   const double *A, int64_t lda, const double *b, double *x,
                   const mqr_ir_opts_t *opts, mqr_ir_info_t *info);

/* ---- batched uneven QR ---------------------------------------------------- */
/* No reference counterpart: the reference factors one MAT per QR() call
 * (QR/QR.h:7-46); a batch there is a loop of such calls.
 * Factor `batch` independent column-major matrices in place with the
 * semantics of mqr_dgeqrf per matrix (R on and above the diagonal, reflectors
 * below, tau[min(m,n)]), in the precision of `dtype`, in one persistent
 * launch.  m, n, lda: host arrays (0 <= m[i] <= 512, n[i] >= 0,
 * lda[i] >= max(1, m[i])); A, tau: host arrays of device pointers (double *
 * for MQR_F64, float * for MQR_F32).  tile: -1 = the tuned per-size table,
 * 0 = panel width 8, 1 = panel width 16.  Blocks until done. */
int mqr_geqrf_batched_uneven(mqr_handle_t handle, mqr_dtype_t dtype, int64_t batch,
                             const int32_t *m, const int32_t *n, const int64_t *lda,
                             void *const *A, void *const *tau, int tile);

/* Allocation-free variant of mqr_geqrf_batched_uneven.  No counterpart in
 * the reference: it is for callers that cannot allocate device memory (code
 * under stream capture, tight serving loops, harnesses that forbid
 * hipMalloc).  Same arguments, results and blocking behaviour; the kernel's
 * descriptor queue and counters live in a fixed device array of the library
 * (about 1.3 MB per GPU, held until the process ends) instead of per-call
 * allocations, and no hipMalloc, hipFree or host heap allocation is made.
 * batch <= MQR_NOALLOC_MAX_BATCH, else MQR_ERR_TOO_LARGE before any work.
 * Every call on one GPU shares that array: calls are serialised within a
 * process, and calls from several processes that overlap on one GPU race. */
#define MQR_NOALLOC_MAX_BATCH 32768
int mqr_geqrf_batched_uneven_noalloc(mqr_handle_t handle, mqr_dtype_t dtype, int64_t batch,
                                     const int32_t *m, const int32_t *n, const int64_t *lda,
                                     void *const *A, void *const *tau, int tile);

/* The same work captured once in a hipGraph: create uploads the descriptors
 * (the device pointers are fixed from then on) and instantiates the graph;
 * run replays it on the handle's stream (stream-ordered, returns at once);
 * destroy frees it. */
typedef struct mqr_batched_plan_st *mqr_batched_plan_t;
int mqr_batched_plan_create(mqr_handle_t handle, mqr_dtype_t dtype, int64_t batch,
                            const int32_t *m, const int32_t *n, const int64_t *lda,
                            void *const *A, void *const *tau, int tile,
                            mqr_batched_plan_t *plan);
/* No reference counterpart: the batched launch's schedule, the space the
 * autotuner (mat_amd/autotune_batched.py) searches.  nb: panel width (0: the
 * tuned width per size bin, else 8 or 16); per_cu: resident workgroups per
 * CU (-1: tuned, 0: as many as the kernel's registers and LDS allow, else
 * at most that many); order: the queue (-1: tuned, 0: decreasing flop
 * count, 1: as given, 2: largest and smallest alternating).  The factors do
 * not depend on the schedule; only which workgroup takes which matrix, and
 * when, does.  A bad schedule is argument -9. */
typedef struct {
    int nb;
    int per_cu;
    int order;
} mqr_batched_schedule_t;
int mqr_batched_plan_create_sched(mqr_handle_t handle, mqr_dtype_t dtype, int64_t batch,
                                  const int32_t *m, const int32_t *n, const int64_t *lda,
                                  void *const *A, void *const *tau,
                                  const mqr_batched_schedule_t *sched, mqr_batched_plan_t *plan);
int mqr_batched_plan_run(mqr_handle_t handle, mqr_batched_plan_t plan);
int mqr_batched_plan_destroy(mqr_batched_plan_t plan);

/* ---- QR adjoint ------------------------------------------------------------ */
/* No reference counterpart.  Reverse-mode derivative of the thin QR A = Q R
 * (m >= n) for `batch` equally sized matrices at fixed strides:
 *   dA = [dQ + Q copyltu(M)] R^{-T},  M = R dR^T - dQ^T Q,
 *   copyltu(M) = tril(M) + tril(M, -1)^T.
 * Q, dQ, dA: m x n; R, dR: n x n (R upper triangular with a nonzero
 * diagonal; only the upper triangle of dR is read, since R's strictly lower
 * part is identically zero).  Matrix i of X is at X + i * strideX (elements).  dQ or dR may
 * be NULL (taken as zero).  All device pointers of `dtype` (double or float);
 * dA must not overlap the inputs. */
int mqr_qr_backward(mqr_handle_t handle, mqr_dtype_t dtype, int64_t batch, int64_t m, int64_t n,
                    const void *Q, int64_t ldq, int64_t strideQ, const void *R, int64_t ldr,
                    int64_t strideR, const void *dQ, int64_t lddq, int64_t stridedQ,
                    const void *dR, int64_t lddr, int64_t stridedR, void *dA, int64_t ldda,
                    int64_t stridedA);

/* ---- host-memory convenience (used by include/mat/Mat.h) ------------------ */
/* QR(A, Q, R, QR_Householder) semantics (QR/QR.h:7-24 with
 * QR/__Householder_Blocked.h:335-353): A m x n (lda = m) is read, Q m x m and
 * R m x n (R zero below the diagonal) are written; all host memory.  Blocks. */
int mqr_qr_householder_host(mqr_handle_t handle, int64_t m, int64_t n, const double *A,
                            double *Q, double *R);
/* QR(A, Q, R, type) on host memory for every QR_type (Mat.h:43-47):
 * A m x n (lda = m) read; R m x n written; Q m x m for MQR_QR_HOUSEHOLDER
 * (R zero below the diagonal) and MQR_QR_GIVENS (R with the rotation
 * residue), m x n for MQR_QR_GRAMSCHMIDT (R zero outside its n x n upper
 * triangle, m >= n).  Blocks. */
typedef enum { MQR_QR_HOUSEHOLDER = 0, MQR_QR_GIVENS = 1, MQR_QR_GRAMSCHMIDT = 2 } mqr_qr_type_t;
int mqr_qr_host(mqr_handle_t handle, mqr_qr_type_t type, int64_t m, int64_t n, const double *A,
                double *Q, double *R);
/* LU(A, L, U) / PLU(A, pos, L, U) on host memory: A m x n (lda = m) read,
 * F (m x n) receives L \ U (L unit lower strictly below the diagonal, U on
 * and above), pos (m) the row permutation (row i of L U is row pos[i] of A;
 * the identity when pivot = 0), info (host, may be NULL) as mqr_dgetrf's.
 * Blocks. */
int mqr_lu_host(mqr_handle_t handle, int pivot, int64_t m, int64_t n, const double *A, double *F,
                int32_t *pos, int32_t *info);
/* X = A^{-1} B (Solve_direct) on host memory: A n x n, B and X n x nrhs.
 * Blocks. */
int mqr_solve_direct_host(mqr_handle_t handle, int64_t n, int64_t nrhs, const double *A,
                          const double *B, double *X);
/* Host least squares: A m x n (lda = m), b (m), x (n) host memory; prec as
 * above (MQR_PREC_FP64 = plain solve, no refinement).  Blocks. */
int mqr_gels_host(mqr_handle_t handle, mqr_precision_t prec, int64_t m, int64_t n,
                  const double *A, const double *b, double *x, mqr_ir_info_t *info);

/* ---- binary MAT files (host memory; no device involved) ------------------ */
/* The reference's file format (Mat-5.2/_extra/__File.h:1-66): int32 rows,
 * int32 cols, then rows * cols elements of elem_size bytes each in the
 * matrix's storage order (column-major for MAT<T> here), no type tag.
 * mqr_mat_load checks that the header equals (rows, cols) and that the file
 * holds exactly that many elements; failures are MQR_ERR_IO.  The raw forms
 * hold elements only (the reference's save / load of a pointer and a length). */
int mqr_file_size(const char *path, int64_t *bytes);                       /* :3-10 */
int mqr_mat_save(const char *path, int32_t rows, int32_t cols, const void *data,
                 size_t elem_size);                                         /* :56-66 */
int mqr_mat_load_header(const char *path, int32_t *rows, int32_t *cols);   /* :21-34 */
int mqr_mat_load(const char *path, int32_t rows, int32_t cols, void *data,
                 size_t elem_size);                                         /* :21-44 */
int mqr_raw_save(const char *path, const void *data, size_t bytes);        /* :48-55 */
int mqr_raw_load(const char *path, void *data, size_t bytes);              /* :12-19 */

#ifdef __cplusplus
}
#endif
#endif /* MAT_QR_H */
