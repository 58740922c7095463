/*
 * mqr_oracle.h -- TEST INFRASTRUCTURE, NOT PRODUCT CODE.
 *
 * CPU restatement of Mat-8.3.1's dense Householder QR path (blocked compact-WY
 * factorisation, packed MR x NR GEMM, unblocked small-size branch, explicit Q),
 * plus a plain least-squares solve on top of it.  Only tests/, the smoke check
 * and bench.py's cpu_baseline leg may load this library, and only as the
 * checker.  The mat_amd product never links, loads or calls it.
 *
 * Storage is column-major everywhere: element (r, c) of a matrix with leading
 * dimension ld sits at data[ld * c + r]  (Base/__Matrix.h:141-146,
 * Base/__Block.h:16-34).
 *
 * The library is compiled twice, once per block-size configuration of the
 * reference (Mat.h:57-109):
 *   AVX2    : MR x NR = 8 x 6,  MC = 576,  KC = 144, NC = 144, BS_QR = 96
 *   AVX-512 : MR x NR = 16 x 12, MC = 1152, KC = 72, NC = 144, BS_QR = 48
 */
#ifndef MQR_ORACLE_H
#define MQR_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Block sizes of this build: out[0..5] = MR, NR, MC, KC, NC, BS_QR. */
void orc_config(int *out);

/* Number of doubles of workspace orc_geqrf_blocked / orc_orgqr_full /
 * orc_larfb_t need for an m x n problem (the arena of QR_WorkAlloc,
 * QR/__Householder_Blocked.h:11-55). */
size_t orc_work_doubles(int m, int n);

/* Level-1 dot with the reference's four-way split accumulation
 * (Calculate/Calculate.h:61-76). */
double orc_dot(const double *a, const double *b, int len);

/* ---- packed GEMM layer (Mul/) ---- */
void orc_pack_A(const double *A, int64_t ldA, int mc, int kc, double *Ap);
void orc_pack_B(const double *B, int64_t ldB, int kc, int nc, double *Bp);
/* C[mm x nn] -= Ap * Bp over k (mul = 0) or C += Ap * Bp (mul = 1) */
void orc_micro_kernel(const double *Ap, const double *Bp, double *C, int64_t ldC,
                      int k, int mm, int nn, int add);
/* C[m x n] -= packedA * packedB  (Mul/Mul.h:239-263) */
void orc_mulAB_packed_n(const double *Ap, const double *Bp, double *C, int64_t ldC,
                        int m, int n, int k);
/* C = A * B, the packed blocked GEMM behind operator* (Mul/Mul.h:154-202). */
void orc_gemm(int M, int N, int K, const double *A, int64_t ldA,
              const double *B, int64_t ldB, double *C, int64_t ldC);

/* ---- Householder QR pieces (QR/__Householder_Blocked.h) ---- */
/* Panel: b Householder reflectors of the h x b panel P, in place.  v is
 * scratch of length h.  (:62-84) */
void orc_panel(int h, int b, double *P, int64_t ldP, double *tau, double *v);
/* Compact-WY T (b x b, ld = b) of the reflectors stored below the diagonal
 * of the h x b panel V.  s, w: scratch of length b.  (:86-103) */
void orc_larft(int h, int b, const double *V, int64_t ldV, const double *tau,
               double *T, double *s, double *w);
/* Trailing update C <- (I - V T V^T)^T C for the h x nc block C.  (:143-212) */
void orc_larfb_t(int h, int b, int nc, const double *V, int64_t ldV,
                 const double *T, double *C, int64_t ldC, double *work);
/* Same reflector block applied as Q: C <- (I - V T V^T) C.  (:276-331) */
void orc_larfb_n(int h, int b, int nc, const double *V, int64_t ldV,
                 const double *T, double *C, int64_t ldC, double *work);
/* Blocked factorisation of the m x n matrix R (ld = m) in place.  T_all holds
 * ceil(k/BS) blocks of BS*BS doubles; block j (first column j) is at
 * T_all + j * BS with leading dimension b_j.  tau has k entries.  (:214-245) */
void orc_geqrf_blocked(int m, int n, double *R, double *T_all, double *tau,
                       double *work);
/* Explicit full Q (m x m, ld = m) from a blocked factorisation.  (:247-333) */
void orc_orgqr_full(int m, int n, const double *VR, const double *T_all,
                    double *Q, double *work);

/* ---- QR/__HouseHolder.h, unblocked small-size branch ---- */
/* R (m x n, ld = m) in place; Q m x m.  (:178-208 with :47-79) */
void orc_qr_unblocked(int m, int n, double *Q, double *R);

/* ---- QR/QR.h:7-46 QR(A, Q, R, QR_Householder) ---- */
/* A m x n (ld = m, read only), Q m x m, R m x n.  Returns 1 when the blocked
 * branch ran, 0 for the unblocked one. */
int orc_qr_householder(int m, int n, const double *A, double *Q, double *R);

/* ---- QR/QR.h:26-42: the other QR_type values ---- */
/* Givens rotations (QR/__Givens.h:45-63): A m x n read, Q m x m and R m x n
 * written (R keeps rotation residue below the diagonal).  fix_zero_pivots
 * selects the GPU path's handling of a zero pivot instead of the source's
 * division by zero. */
void orc_qr_givens(int m, int n, const double *A, double *Q, double *R, int fix_zero_pivots);
/* Modified Gram-Schmidt (QR/__GramSchmidt.h:2-18): A m x n read, Q m x n
 * and R m x n (upper n x n triangle, zero elsewhere) written. */
void orc_qr_gramschmidt(int m, int n, const double *A, double *Q, double *R);

/* ---- LU / PLU / Solve_direct (LU/__LU.h, LU/__PLU.h, LU/__Solve_direct.h) ---- */
/* In place on F (m x n, ld m): L unit lower (strictly below the diagonal),
 * U on and above; pivot = 1 for partial pivoting (first row of largest
 * magnitude), pos[m] the row permutation (row i of L U is row pos[i] of A).
 * Returns the number of columns factored (min(m, n) unless a zero pivot /
 * zero column stopped it). */
int orc_lu_factor(int m, int n, double *F, int pivot, int *pos);
/* X = A^{-1} B by PLU, A n x n, B and X n x nrhs (ld n). */
void orc_solve_direct(int n, int nrhs, const double *A, const double *B, double *X);

/* ---- least squares (no reference counterpart) ---- */
/* x = argmin ||A x - b|| for m >= n via the blocked factorisation, Q^T b
 * through the compact-WY blocks, and back substitution with R
 * (pattern of LU/__Solve_direct.h:39-65).  A and b are not modified. */
void orc_gels(int m, int n, const double *A, const double *b, double *x);

/* Back substitution R x = c for upper-triangular R (n x n, ld = ldR). */
void orc_trsv_upper(int n, const double *R, int64_t ldR, double *x);

#ifdef __cplusplus
}
#endif
#endif
