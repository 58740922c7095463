// Latency-optimal panel factorisation for tall panels.
//
// The reference factors each h x b panel column by column
// (QR/__Householder_Blocked.h:62-84): b dependent reductions over h rows.  On
// the GPU every such reduction is a grid-wide synchronisation, so the tall
// panel is factored instead by CholeskyQR2: two passes of
//   Gram (MFMA, split-K) -> Cholesky (+ triangular inverse) -> Y = X R^{-1}
// (orthonormal to working accuracy when kappa(X) is below ~1e6).  Every
// workgroup of the kernel forming the first Y repeats the first Cholesky;
// the second runs in the single-workgroup reconstruction.  The
// Householder vectors are then reconstructed from the orthonormal factor
// Q = Y R2^{-1} by an LU factorisation of Q(0:b, :) - S without pivoting,
// S = diag(+-1) chosen during the elimination (S_j = -sign of the current
// diagonal, sign(0) = +):  V1 = L, V2 = Q(b:h, :) U^{-1}, tau_j = -S_j U_jj,
// T = -U S L^{-T}, R = S R2 R1.  In exact arithmetic these are the
// reference's reflectors and R (R_jj = -sign(alpha_j) |R_jj|; a column with
// nothing left below the diagonal is reflected with tau = 2).
//
// A first-pass pivot d_i <= 1e-12 G_ii (a zero, dependent or nearly
// dependent column), a second-pass pivot below 1e-2, or a non-finite value
// raises a device flag; the rest of the sequence then leaves the panel
// untouched and the reconstruction kernel factors it column by column
// instead (panel_hh.h), which also keeps the reference's tau = 0 for zero
// columns.
//
// The mixed-precision factorisation, whose FP32 trailing updates perturb the
// factors by about 6e-8 anyway, lets the first pass decide on the device to
// skip the second when the first Cholesky factor predicts Z orthonormal to
// 1e-9 (k_upd_apply); the FP64 factorisation always takes both passes.
//
// The b x b factorisations run register-resident in one workgroup
// (small_dense.h); every barrier is executed by all of its waves.  The
// compact-WY factor T comes out of the reconstruction as well.
#include <algorithm>
#include <climits>

#include "mqr_internal.h"
#include "panel_hh.h"
#include "small_blocked.h"
#include "small_dense.h"

namespace mqr {
namespace {

constexpr int PB = kMaxNb;  // padded panel width (small::N)
constexpr int kFbCounters = 128;  // barrier counters of the breakdown fallback

// Second pass and reconstruction.  Inputs: reduced Gram G2 of Z = X R1^{-1}
// (b x b, ld b), the top b rows of Z (ld ldz), R1 (padded, ld PB).  Outputs:
// the panel's top b x b block (R = S R2 R1 on/above the diagonal, V1 below),
// tau, the compact-WY factor T (ld ldt) and Minv = R2^{-1} U^{-1} (padded)
// for V2 = Z(b:h) Minv.
//
// With Q = Z R2^{-1}, Q1 = Q(0:b, :), and the LU factorisation
// Q1 - S = L U (S the signs), the reflectors I - V T V^T whose first b
// columns are Q S satisfy V1 = L, V2 = Q2 U^{-1} and T = -U S L^{-T}
// (top block: Q1 S - I = -V1 T V1^T).
//
// Two 64 x 64 matrices of LDS (67 KB), so the kernel starts on a CU that
// still runs one FP64 update tile (66.5 KB) instead of waiting for an empty
// one.  Operands that only one wave reads as its tile row stay in that
// wave's registers (Z's top rows, R2) or, for R2^{-1} between its inverse
// and Minv, in the 64 x 64 global scratch `stash` (each lane re-reads the
// values it wrote).
//
// Breakdown: when the first pass raised the flag, or the second Cholesky
// fails here, the untouched h x b panel A is factored column by column in
// this same workgroup (the reference's panel loop), with the flag left up
// for the V2 kernel; no separate fallback launch.
//
// With one_pass and *one_pass set (by k_upd_apply, when its first factor
// predicts Z orthonormal to 1e-9), Z itself is the orthonormal factor:
// R2 = I, so Q1 = Z_top, R = S R1 and Minv = U^{-1}, and G is not read.
__global__ __launch_bounds__(small::NT) void k_recon(int64_t h, int b,
                                                     const double *__restrict__ G, int64_t ldg,
                                                     const double *__restrict__ Z, int64_t ldz,
                                                     const double *__restrict__ R1,
                                                     double *__restrict__ A, int64_t lda,
                                                     double *__restrict__ tau,
                                                     double *__restrict__ T, int64_t ldt,
                                                     double *__restrict__ Minv,
                                                     double *__restrict__ stash,
                                                     int *__restrict__ flag,
                                                     const int *__restrict__ one_pass) {
    __shared__ small::Lds P[PB], Q[PB];
    __shared__ double sg[PB], red[small::NT / 64];
    __shared__ int fail;
    // the fallback's LDS: P, and its small buffers carved from Q
    double *const qa = &Q[0][0];
    const hh::PanelLds fb{reinterpret_cast<double(*)[hh::CH]>(qa), qa + 64, P,
                          reinterpret_cast<small::RowBuf *>(qa + 128), qa + 128 + 2 * 4 * PB};
    static_assert(128 + 2 * 4 * PB + PB <= PB * (PB + 1), "fallback LDS exceeds Q");
    static_assert(sizeof(small::RowBuf) == 2 * 4 * PB * sizeof(double), "RowBuf layout");
    const int f0 = *flag;
    if (f0 == 2) return;  // factored by k_apply_gram's workgroups
    if (f0) {
        hh::panel_hh_body(h, b, A, lda, tau, T, ldt, fb);
        return;
    }
    const int tid = threadIdx.x, q = small::wave_id();
    const int l = small::lane_id(), li = l & 15, kq = l >> 4, arow = 16 * q + li;
    sb::v4d rr[4];  // tile row q of R2 R1
    if (one_pass && *one_pass) {
        // Q1 = Z_top -> P (padded with I), R2 R1 = R1
#pragma unroll
        for (int jb = 0; jb < 4; ++jb)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int i = 16 * q + sb::trow(r), c = 16 * jb + sb::tcol();
                P[i][c] = (i < b && c < b) ? Z[i + (int64_t)c * ldz] : (i == c ? 1.0 : 0.0);
                rr[jb][r] = R1[i + (int64_t)c * PB];
            }
        __syncthreads();
        // LU of Q1 - S -> P; U^{-1} -> Q, which is Minv
        sb::lu_sign_invu(P, Q, sg);
#pragma unroll
        for (int jb = 0; jb < 4; ++jb)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int i = 16 * q + sb::trow(r), c = 16 * jb + sb::tcol();
                Minv[i + (int64_t)c * PB] = i <= c ? Q[i][c] : 0.0;
            }
        __syncthreads();  // U^{-1} no longer read
    } else {
        if (tid == 0) fail = 0;
        small::lds_load(P, G, ldg, b);
        __syncthreads();
        // e2 = ||G2 - I||_F^2, the same in every wave
        double e2 = 0.0;
        for (int e = tid; e < PB * PB; e += small::NT) {
            const int i = e / PB, j = e % PB;
            const double d = P[i][j] - (i == j ? 1.0 : 0.0);
            e2 = fma(d, d, e2);
        }
        e2 = small::wave_sum(e2);
        if (l == 0) red[q] = e2;
        __syncthreads();
        e2 = (red[0] + red[1]) + (red[2] + red[3]);
        if (e2 <= 1e-12) {
            // G2 = I + E with |E|_F <= 1e-6: R2 = I + F and R2^{-1} from
            // series instead of the Cholesky's 64 dependent pivots.  With U(X)
            // the strict upper triangle of X plus half its diagonal,
            // F = U(E) - U(U(E)^T U(E)) satisfies F + F^T + F^T F = E up to
            // O(|E|^3), and R2^{-1} = I - F + F^2 up to O(|E|^3) (<= 1e-18).
            // U(E) -> Q
            for (int e = tid; e < PB * PB; e += small::NT) {
                const int i = e / PB, j = e % PB;
                Q[i][j] = j > i ? P[i][j] : (j == i ? 0.5 * (P[i][i] - 1.0) : 0.0);
            }
            __syncthreads();
            // F -> P, tile row q (U(E) is upper: k-tiles 0..q of U(E)^T U(E))
#pragma unroll
            for (int jb = 0; jb < 4; ++jb) {
                sb::v4d sq = sb::zero4();
                if (jb >= q)
                    for (int k = 0; k <= q; ++k)
                        sq = sb::tmma<true, false>(sq, Q, q, k, Q, k, jb);
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int i = 16 * q + sb::trow(r), j = 16 * jb + sb::tcol();
                    P[i][j] = j > i ? Q[i][j] - sq[r] : (j == i ? Q[i][i] - 0.5 * sq[r] : 0.0);
                }
            }
            __syncthreads();
            // R2^{-1} = I - F + F F -> Q, tile row q
#pragma unroll
            for (int jb = 0; jb < 4; ++jb) {
                sb::v4d f2 = sb::zero4();
                if (jb >= q)
                    for (int k = q; k <= jb; ++k) f2 = sb::tmma<false, false>(f2, P, q, k, P, k, jb);
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int i = 16 * q + sb::trow(r), j = 16 * jb + sb::tcol();
                    Q[i][j] = j >= i ? (i == j ? 1.0 : 0.0) - P[i][j] + f2[r] : 0.0;
                }
            }
            __syncthreads();
            // R2 = I + F
            for (int i = tid; i < PB; i += small::NT) P[i][i] += 1.0;
            __syncthreads();
        } else {
            // R2 = chol(G2) -> P, R2^{-1} -> Q
            sb::chol_inv(P, Q, b, [](int) { return 1e-2; }, &fail);
        }
        if (fail) {  // uniform: read after the barrier that ends chol_inv
            if (tid == 0) *flag = 1;
            __syncthreads();  // the LDS buffers are reused by the fallback
            hh::panel_hh_body(h, b, A, lda, tau, T, ldt, fb);
            return;
        }
        // tile row q of R2 as MFMA operands; R2^{-1}'s to the stash
        double r2[4][4];
#pragma unroll
        for (int k = 0; k < 4; ++k)
#pragma unroll
            for (int k4 = 0; k4 < 4; ++k4) {
                const int c = 16 * k + 4 * k4 + kq;
                r2[k][k4] = P[arow][c];
                stash[arow + (int64_t)c * PB] = Q[arow][c];
            }
        // Q1 = Z_top R2^{-1} -> tile row q of P (Z_top padded with I like G2)
        {
            double z[4][4];
#pragma unroll
            for (int k = 0; k < 4; ++k)
#pragma unroll
                for (int k4 = 0; k4 < 4; ++k4) {
                    const int c = 16 * k + 4 * k4 + kq;
                    z[k][k4] = (arow < b && c < b) ? Z[arow + (int64_t)c * ldz] : (arow == c ? 1.0 : 0.0);
                }
            sb::v4d acc[4];
#pragma unroll
            for (int jb = 0; jb < 4; ++jb) {
                acc[jb] = sb::zero4();
#pragma unroll
                for (int k = 0; k < 4; ++k)
                    if (k <= jb) acc[jb] = sb::tmma_reg(acc[jb], z[k], Q, k, jb);
            }
            // this wave's reads of its tile row of P are done before it writes it
            __builtin_amdgcn_wave_barrier();
#pragma unroll
            for (int jb = 0; jb < 4; ++jb) sb::tstore(P, q, jb, acc[jb]);
        }
        __syncthreads();  // R2^{-1} no longer read
        // RR = R2 R1, kept in registers (tile row q)
        small::lds_load(Q, R1, PB, PB);
        __syncthreads();
#pragma unroll
        for (int jb = 0; jb < 4; ++jb) {
            rr[jb] = sb::zero4();
#pragma unroll
            for (int k = 0; k < 4; ++k)
                if (k >= q && k <= jb) rr[jb] = sb::tmma_reg(rr[jb], r2[k], Q, k, jb);
        }
        __syncthreads();  // R1 no longer read
        // LU of Q1 - S -> P; U^{-1} -> Q
        sb::lu_sign_invu(P, Q, sg);
        // Minv = R2^{-1} U^{-1} straight to global, R2^{-1} back from the stash
        {
            double ri[4][4];
#pragma unroll
            for (int k = 0; k < 4; ++k)
#pragma unroll
                for (int k4 = 0; k4 < 4; ++k4)
                    ri[k][k4] = stash[arow + (int64_t)(16 * k + 4 * k4 + kq) * PB];
#pragma unroll
            for (int jb = 0; jb < 4; ++jb) {
                sb::v4d acc = sb::zero4();
#pragma unroll
                for (int k = 0; k < 4; ++k)
                    if (k >= q && k <= jb) acc = sb::tmma_reg(acc, ri[k], Q, k, jb);
#pragma unroll
                for (int r = 0; r < 4; ++r)
                    Minv[16 * q + sb::trow(r) + (int64_t)(16 * jb + sb::tcol()) * PB] = acc[r];
            }
        }
        __syncthreads();  // U^{-1} no longer read
    }
    // L^{-1} -> Q
    sb::inv_unit_lower(P, Q);
    // T = -(U S) L^{-T} (upper), R = S (R2 R1) with V1 strictly below
#pragma unroll
    for (int jb = 0; jb < 4; ++jb) {
        sb::v4d acc = sb::zero4();
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            if (k < q || k > jb) continue;
#pragma unroll
            for (int k4 = 0; k4 < 4; ++k4) {
                const int c = 16 * k + 4 * k4 + kq;
                const double us = arow <= c ? P[arow][c] * sg[c] : 0.0;
                acc = __builtin_amdgcn_mfma_f64_16x16x4f64(-1.0 * us, Q[16 * jb + li][c], acc, 0,
                                                           0, 0);
            }
        }
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int rr_ = 16 * q + sb::trow(r), cc = 16 * jb + sb::tcol();
            if (rr_ < b && cc < b) {
                T[rr_ + (int64_t)cc * ldt] = rr_ <= cc ? acc[r] : 0.0;
                A[rr_ + (int64_t)cc * lda] = rr_ <= cc ? sg[rr_] * rr[jb][r] : P[rr_][cc];
            }
        }
    }
    for (int i = tid; i < b; i += small::NT) tau[i] = -sg[i] * P[i][i];
}

// ---- row-chunk kernels of the tall-panel sequence ----------------------
//
// Each workgroup streams 64-row chunks of the h x b panel through LDS, forms
// its products on the FP64 MFMA and, where a Gram matrix is needed, adds the
// chunk's contribution to the upper 64 x 64 tiles it owns; one partial per
// workgroup is summed by reduce_partials.  Results are held in the
// row-per-lane tile layout (lane l, register r: row 16 i + (l & 15), column
// 16 j + (l >> 4) + 4 r), so global loads and stores of a 64-row chunk are
// contiguous 128-byte column pieces.

constexpr int RC = 64;              // rows per chunk
constexpr int kMaxGramParts = 256;  // Gram partials per pass (workgroups)

// acc += s op(A)[ia, ka] op(B)[kb, jb] in the row-per-lane layout
template <bool TA, bool TB>
__device__ __forceinline__ sb::v4d tmmaT(sb::v4d acc, const small::Lds *A, int ia, int ka,
                                         const small::Lds *B, int kb, int jb, double s = 1.0) {
    const int l = small::lane_id(), i = l & 15, kq = l >> 4;
#pragma unroll
    for (int k4 = 0; k4 < 4; ++k4) {
        const int ak = 16 * ka + 4 * k4 + kq, bk = 16 * kb + 4 * k4 + kq;
        const double av = TA ? A[ak][16 * ia + i] : A[16 * ia + i][ak];
        const double bv = TB ? B[16 * jb + i][bk] : B[bk][16 * jb + i];
        acc = __builtin_amdgcn_mfma_f64_16x16x4f64(bv, s * av, acc, 0, 0, 0);
    }
    return acc;
}

// acc[jb] = X[tile row q] M[:, tile jb] for the upper-triangular M (k-blocks
// <= jb), the four column tiles interleaved.
__device__ __forceinline__ void rowtile_upper(sb::v4d (&acc)[4], const small::Lds *X,
                                              const small::Lds *M, int q) {
    const int l = small::lane_id(), i = l & 15, kq = l >> 4;
#pragma unroll
    for (int jb = 0; jb < 4; ++jb) acc[jb] = sb::zero4();
#pragma unroll
    for (int k = 0; k < 16; ++k) {
        const int kk = 4 * k + kq;
        const double av = X[16 * q + i][kk];
#pragma unroll
        for (int jb = 0; jb < 4; ++jb)
            if (jb >= (k >> 2))
                acc[jb] = __builtin_amdgcn_mfma_f64_16x16x4f64(M[kk][16 * jb + i], av, acc[jb], 0,
                                                               0, 0);
    }
}

__device__ __forceinline__ int lrow() { return small::lane_id() & 15; }
__device__ __forceinline__ int lcol(int r) { return (small::lane_id() >> 4) + 4 * r; }

__device__ __forceinline__ void tstoreT(small::Lds *S, int ia, int jb, sb::v4d v) {
#pragma unroll
    for (int r = 0; r < 4; ++r) S[16 * ia + lrow()][16 * jb + lcol(r)] = v[r];
}

// v = the calling thread's values of M[r0 + r, c] (r0 + r < rows and
// c < cols, else 0) in the pair layout (small::pair_fetch)
template <typename TM>
__device__ __forceinline__ void chunk_fetch(double (&v)[small::TV], const TM *__restrict__ M,
                                            int64_t ld, int64_t r0, int64_t rows, int cols) {
    static_assert(RC == small::N && PB == small::N, "chunk shape");
    small::pair_fetch(v, M + r0, ld, rows - r0, cols);
}

// S[r][c] = M[r0 + r, c] for r0 + r < rows and c < cols, else 0
template <typename TM>
__device__ __forceinline__ void chunk_load(small::Lds *S, const TM *__restrict__ M, int64_t ld,
                                           int64_t r0, int64_t rows, int cols) {
    double v[small::TV];
    chunk_fetch(v, M, ld, r0, rows, cols);
    small::pair_stash(S, v);
}

// upper tile u (0..9) in row order: (0,0) (0,1) (0,2) (0,3) (1,1) ... (3,3)
__device__ __forceinline__ void upper_tile(int u, int &i, int &j) {
    if (u < 4) { i = 0; j = u; }
    else if (u < 7) { i = 1; j = u - 3; }
    else if (u < 9) { i = 2; j = u - 5; }
    else { i = 3; j = 3; }
}
// strictly lower tile u (0..5): (1,0) (2,0) (2,1) (3,0) (3,1) (3,2)
__device__ __forceinline__ void lower_tile(int u, int &i, int &j) {
    if (u < 1) { i = 1; j = 0; }
    else if (u < 3) { i = 2; j = u - 1; }
    else { i = 3; j = u - 3; }
}

// g += S^T S on the upper tiles wave q owns (tiles q, q + 4, q + 8); the
// tiles' MFMAs are interleaved so that independent accumulators pipeline.
__device__ __forceinline__ void gram_acc(const small::Lds *S, sb::v4d (&g)[3]) {
    const int q = small::wave_id(), l = small::lane_id(), i = l & 15, kq = l >> 4;
    int ti[3], tj[3];
#pragma unroll
    for (int u = 0; u < 3; ++u) upper_tile(q + 4 * u < 10 ? q + 4 * u : 9, ti[u], tj[u]);
    const bool third = q + 8 < 10;
#pragma unroll
    for (int k = 0; k < 16; ++k) {  // 4-row slices of the 64-row chunk
        const int kk = 4 * k + kq;
#pragma unroll
        for (int u = 0; u < 3; ++u) {
            if (u == 2 && !third) continue;
            const double av = S[kk][16 * ti[u] + i], bv = S[kk][16 * tj[u] + i];
            g[u] = __builtin_amdgcn_mfma_f64_16x16x4f64(bv, av, g[u], 0, 0, 0);
        }
    }
}

// Gp (64 x 64, ld 64) = the upper tiles in g, zero below the diagonal tiles
__device__ __forceinline__ void gram_store(const sb::v4d (&g)[3], double *__restrict__ Gp) {
    const int q = small::wave_id();
#pragma unroll
    for (int u = 0; u < 3; ++u) {
        const int t = q + 4 * u;
        if (t < 10) {
            int i, j;
            upper_tile(t, i, j);
#pragma unroll
            for (int r = 0; r < 4; ++r) Gp[16 * i + lrow() + (16 * j + lcol(r)) * PB] = g[u][r];
        }
        if (t < 6) {
            int i, j;
            lower_tile(t, i, j);
#pragma unroll
            for (int r = 0; r < 4; ++r) Gp[16 * i + lrow() + (16 * j + lcol(r)) * PB] = 0.0;
        }
    }
}

// C (hc x bn, ldc, FP64 or FP32) -= Y (hc x kin, ldy) W (kin x bn, ldw)
// when kin > 0 (in FP64), then the Gram partial of rows [g0, hc) of C into
// Gp[blockIdx.x]; with Xo, also an FP64 copy of those rows into
// Xo[row - g0 + col * ldxo] (the panel's FP64 working copy when C is FP32).
template <typename TC>
__global__ __launch_bounds__(small::NT) void k_upd_gram(int64_t hc, int kin, int bn,
                                                        const double *__restrict__ Y, int64_t ldy,
                                                        const double *__restrict__ W, int64_t ldw,
                                                        TC *__restrict__ C, int64_t ldc,
                                                        int64_t g0, double *__restrict__ Gp,
                                                        double *__restrict__ Xo, int64_t ldxo,
                                                        int *__restrict__ bar) {
    __shared__ small::Lds Ys[RC], Ws[PB];
    const int q = small::wave_id();
    // the next k_apply_gram's fallback barriers start from zero
    if (blockIdx.x == 0 && threadIdx.x < kFbCounters) bar[threadIdx.x] = 0;
    sb::v4d g[3] = {sb::zero4(), sb::zero4(), sb::zero4()};
    if (kin > 0) {
        double wv[small::TV];
        small::pair_fetch(wv, W, ldw, kin, bn);
        small::pair_stash(Ws, wv);
    }
    for (int64_t r0 = (int64_t)blockIdx.x * RC; r0 < hc; r0 += (int64_t)gridDim.x * RC) {
        __syncthreads();  // the previous chunk's Gram reads are done
        if (kin > 0) chunk_load(Ys, Y, ldy, r0, hc, kin);
        const int64_t row = r0 + 16 * q + lrow();
        sb::v4d acc[4];
#pragma unroll
        for (int jb = 0; jb < 4; ++jb)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int col = 16 * jb + lcol(r);
                acc[jb][r] = (row < hc && col < bn) ? (double)C[row + (int64_t)col * ldc] : 0.0;
            }
        __syncthreads();
        if (kin > 0) {
            const int l = small::lane_id(), i = l & 15, kq = l >> 4;
            for (int k = 0; 4 * k < kin; ++k) {
                const int kk = 4 * k + kq;
                const double av = -Ys[16 * q + i][kk];
#pragma unroll
                for (int jb = 0; jb < 4; ++jb)
                    acc[jb] = __builtin_amdgcn_mfma_f64_16x16x4f64(Ws[kk][16 * jb + i], av,
                                                                   acc[jb], 0, 0, 0);
            }
#pragma unroll
            for (int jb = 0; jb < 4; ++jb)
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int col = 16 * jb + lcol(r);
                    if (row < hc && col < bn) C[row + (int64_t)col * ldc] = (TC)acc[jb][r];
                }
        }
        if (Xo) {
#pragma unroll
            for (int jb = 0; jb < 4; ++jb)
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int col = 16 * jb + lcol(r);
                    if (row >= g0 && row < hc && col < bn)
                        Xo[row - g0 + (int64_t)col * ldxo] = acc[jb][r];
                }
        }
        __syncthreads();  // Ys no longer read by the update
#pragma unroll
        for (int jb = 0; jb < 4; ++jb) {
            sb::v4d v = acc[jb];
            if (row < g0) v = sb::zero4();
            tstoreT(Ys, q, jb, v);
        }
        __syncthreads();
        gram_acc(Ys, g);
    }
    gram_store(g, Gp + (size_t)blockIdx.x * PB * PB);
}

// ---- breakdown fallback on all of k_apply_gram's workgroups --------------
//
// When the first Cholesky fails, every workgroup of k_apply_gram knows it
// (they all factored the same G1), and together they factor the panel by
// Householder reflections column by column, as the reference does
// (QR/__Householder_Blocked.h:62-84), instead of leaving it to one
// workgroup.  Each workgroup keeps its row chunks (at most two of 64 rows)
// in LDS; per column it publishes its partial sums (||x||^2 below the
// diagonal and x^T a_c for the later columns) and, from the chunk holding
// the diagonal row, the top row, then passes a grid barrier, sums the
// partials of all workgroups in a fixed order (so every workgroup derives
// the same tau and coefficients), and updates its rows.  T then comes from
// G = V^T V: T^{-1} = strict_upper(G) + diag(1 / tau).
//
// The barriers are bounded: a workgroup that waits longer than 20 ms (its
// peers could not all become resident beside the other streams' work)
// cancels the barrier, and every workgroup leaves.  The panel X is only
// written after the last barrier, so a cancelled attempt leaves it as it
// was, with the flag at 1 for the single-workgroup fallback in k_recon;
// a completed one sets the flag to 2.
constexpr int kBarCancelled = INT_MIN / 2;
constexpr uint64_t kBarTimeoutTicks = 2000000;  // 20 ms of the 100 MHz real-time counter
constexpr int kFbMaxChunks = 2;                  // row chunks per workgroup held in LDS

// Grid barrier on counter *cnt (zero before first use) for nwg workgroups;
// false when it was cancelled.  Every thread of the workgroup calls it.
__device__ bool grid_sync(int *cnt, int nwg, int *ok) {
    __syncthreads();
    if (threadIdx.x == 0) {
        __threadfence();
        int v = atomicAdd(cnt, 1) + 1;
        bool res = v > 0;
        if (res && v < nwg) {
            const uint64_t t0 = __builtin_amdgcn_s_memrealtime();
            for (;;) {
                v = __hip_atomic_load(cnt, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                if (v >= nwg) break;
                if (v < 0) {
                    res = false;
                    break;
                }
                if (__builtin_amdgcn_s_memrealtime() - t0 > kBarTimeoutTicks) {
                    // cancel unless the last workgroup arrived meanwhile
                    if (atomicCAS(cnt, v, kBarCancelled) == v) {
                        res = false;
                        break;
                    }
                    continue;
                }
                __builtin_amdgcn_s_sleep(2);
            }
        }
        __threadfence();
        *ok = res ? 1 : 0;
    }
    __syncthreads();
    return *ok != 0;
}

// g += V^T V over one LDS chunk of rows r0.. (h rows in all): V(i, c) =
// S[i - r0][c] below the diagonal, 1 on it, 0 above and past row h.
__device__ __forceinline__ void gram_acc_v(const small::Lds *S, int64_t r0, int64_t h, int b,
                                           sb::v4d (&g)[3]) {
    const int q = small::wave_id(), l = small::lane_id(), i = l & 15, kq = l >> 4;
    int ti[3], tj[3];
#pragma unroll
    for (int u = 0; u < 3; ++u) upper_tile(q + 4 * u < 10 ? q + 4 * u : 9, ti[u], tj[u]);
    const bool third = q + 8 < 10;
    auto v = [&](int kk, int c) -> double {
        const int64_t gr = r0 + kk;
        if (gr >= h || c >= b) return 0.0;
        return gr > c ? S[kk][c] : (gr == c ? 1.0 : 0.0);
    };
#pragma unroll
    for (int k = 0; k < 16; ++k) {
        const int kk = 4 * k + kq;
#pragma unroll
        for (int u = 0; u < 3; ++u) {
            if (u == 2 && !third) continue;
            const double av = v(kk, 16 * ti[u] + i), bv = v(kk, 16 * tj[u] + i);
            g[u] = __builtin_amdgcn_mfma_f64_16x16x4f64(bv, av, g[u], 0, 0, 0);
        }
    }
}

struct FbLds {
    double dsum[PB], tops[PB], cfs[PB], tauv[PB];
    double quarter[4][PB];
    int ok;
};

// The grid fallback (see above).  S[0..nch) hold this workgroup's chunks;
// Gp (partials), G (64 x 64, ld PB) and bar come from the panel workspace.
__device__ __forceinline__ void fallback_grid(int64_t h, int b, double *__restrict__ X, int64_t ldx,
                              double *__restrict__ Z, double *__restrict__ Gp,
                              double *__restrict__ G,
                              double *__restrict__ tau, double *__restrict__ T, int64_t ldt,
                              int *__restrict__ flag, int *__restrict__ bar, small::Lds *S0,
                              small::Lds *S1, FbLds &L) {
    const int tid = threadIdx.x, lr = tid & 63, cg = tid >> 6, lane = small::lane_id();
    // as few workgroups as hold every chunk in LDS (each one reads all
    // workgroups' partials per column, so their number counts twice)
    const int chunks = (int)((h + RC - 1) / RC);
    const int nwg = (chunks + kFbMaxChunks - 1) / kFbMaxChunks, wg = blockIdx.x;
    if (wg >= nwg) return;  // uniform; before any barrier
    small::Lds *S[kFbMaxChunks] = {S0, S1};
    int64_t r0[kFbMaxChunks];
    int nch = 0;
    for (int k = 0; k < kFbMaxChunks; ++k) {
        r0[k] = ((int64_t)wg + (int64_t)k * nwg) * RC;
        if (r0[k] < h) nch = k + 1;
    }
    __syncthreads();  // the caller's last reads of the LDS buffers are done
    for (int k = 0; k < nch; ++k) chunk_load(S[k], X, ldx, r0[k], h, b);
    // per-column partials and top rows in Z (unused by the fallback; it
    // holds h x PB >= 2 nwg PB + 2 PB doubles), the Gram partials in Gp
    double *part = Z;                          // [2][nwg][PB]
    double *top = Z + (size_t)2 * nwg * PB;   // [2][PB]
    int nbar = 0;
    // partial sums of column r (rows below r) into slot r & 1, and the top
    // row from the chunk holding row r
    auto publish = [&](int r) {
        __syncthreads();
        for (int ch = 0; ch < nch; ++ch)
            if (r0[ch] + lr == r)
                for (int c = r + cg; c < b; c += 4) top[(r & 1) * PB + c] = S[ch][lr][c];
        // column c = cg + 4k of wave cg: sum over the rows below r
        for (int c = cg + 4 * ((r - cg + 3) >> 2 > 0 ? (r - cg + 3) >> 2 : 0); c < b; c += 4) {
            double a = 0.0;
            for (int ch = 0; ch < nch; ++ch) {
                const int64_t i = r0[ch] + lr;
                if (i > r && i < h) a = fma(S[ch][lr][r], S[ch][lr][c], a);
            }
            const double sum = small::wave_sum(a);  // c is uniform in the wave
            if (lane == 0) part[((size_t)(r & 1) * nwg + wg) * PB + c] = sum;
        }
    };
    publish(0);
    for (int r = 0; r < b; ++r) {
        if (!grid_sync(bar + nbar++, nwg, &L.ok)) return;
        // column sums in a fixed order (each wave a quarter of the
        // workgroups, then the four quarters), the top row
        {
            const int c = lr, w0q = (cg * nwg) >> 2, w1q = ((cg + 1) * nwg) >> 2;
            double sum = 0.0;
            if (c >= r && c < b)
                for (int w = w0q; w < w1q; ++w) sum += part[((size_t)(r & 1) * nwg + w) * PB + c];
            L.quarter[cg][c] = sum;
        }
        __syncthreads();
        if (tid < PB) {
            const int c = tid;
            L.dsum[c] = (L.quarter[0][c] + L.quarter[1][c]) + (L.quarter[2][c] + L.quarter[3][c]);
            L.tops[c] = (c >= r && c < b) ? top[(r & 1) * PB + c] : 0.0;
        }
        __syncthreads();
        const double alpha = L.tops[r], nrm = sqrt(alpha * alpha + L.dsum[r]);
        const bool zero = nrm == 0.0;
        const double cr = alpha < 0.0 ? -nrm : nrm;
        const double w0 = alpha + cr;
        const double hr = zero ? 0.0 : 1.0 / w0, t = zero ? 0.0 : w0 / cr;
        if (tid < PB) {
            const int c = tid;
            L.cfs[c] = (c > r && c < b) ? -t * (L.tops[c] + hr * L.dsum[c]) : 0.0;
            if (c == r) L.tauv[r] = t;
        }
        __syncthreads();
        if (!zero) {
            // v_i = x_i hr below the diagonal, 1 on it; a_c += cf_c v
            for (int ch = 0; ch < nch; ++ch) {
                const int64_t i = r0[ch] + lr;
                if (i < r || i >= h) continue;
                const double x = S[ch][lr][r];
                const double v = i == r ? 1.0 : x * hr;
                for (int c = r + 1 + ((cg - (r + 1)) & 3); c < b; c += 4)
                    S[ch][lr][c] = fma(L.cfs[c], v, S[ch][lr][c]);
            }
            __syncthreads();  // column r read by every wave above
            if (cg == (r & 3))
                for (int ch = 0; ch < nch; ++ch) {
                    const int64_t i = r0[ch] + lr;
                    if (i == r) S[ch][lr][r] = -cr;
                    else if (i > r && i < h) S[ch][lr][r] *= hr;
                }
        }
        if (r + 1 < b) publish(r + 1);
    }
    // G = V^T V: partials, then the 4096 sums split over the workgroups
    __syncthreads();
    sb::v4d g[3] = {sb::zero4(), sb::zero4(), sb::zero4()};
    for (int ch = 0; ch < nch; ++ch) gram_acc_v(S[ch], r0[ch], h, b, g);
    gram_store(g, Gp + (size_t)wg * PB * PB);
    if (!grid_sync(bar + nbar++, nwg, &L.ok)) return;
    for (int e = wg * small::NT + tid; e < PB * PB; e += nwg * small::NT) {
        double sum = 0.0;
        for (int w = 0; w < nwg; ++w) sum += Gp[(size_t)w * PB * PB + e];
        G[e] = sum;
    }
    if (!grid_sync(bar + nbar++, nwg, &L.ok)) return;
    // no barrier can be cancelled any more: write the factored rows back
    for (int ch = 0; ch < nch; ++ch)
        for (int e = tid; e < RC * PB; e += small::NT) {
            const int rr = e & (RC - 1), c = e >> 6;
            if (r0[ch] + rr < h && c < b) X[r0[ch] + rr + (int64_t)c * ldx] = S[ch][rr][c];
        }
    if (wg != 0) return;
    // T (b x b) by T(i, c) = tau_i (delta_ic - sum_{i < l <= c} G(i, l) T(l, c)),
    // one column per thread, in LDS (S0 is free: its rows are written back)
    __syncthreads();
    for (int e = tid; e < PB * PB; e += small::NT) S0[e & (PB - 1)][e >> 6] = G[e];
    __syncthreads();
    small::Lds *Tl = S1;  // S1 is free as well
    if (tid < b) {
        const int c = tid;
        for (int i = c; i >= 0; --i) {
            double sum = 0.0;
            for (int l = i + 1; l <= c; ++l) sum = fma(S0[i][l], Tl[l][c], sum);
            Tl[i][c] = L.tauv[i] * ((i == c ? 1.0 : 0.0) - sum);
        }
        for (int i = 0; i < b; ++i) T[i + (int64_t)c * ldt] = i <= c ? Tl[i][c] : 0.0;
        tau[c] = L.tauv[c];
    }
    if (tid == 0) *flag = 2;
}

// First-pass Cholesky and second-pass Gram.  Every workgroup factors the
// reduced Gram matrix G1 (b x b, ld PB) = R1^T R1 and inverts R1 itself:
// the same operations on the same data, so all workgroups hold the same R1
// bit for bit and no single-workgroup launch sits between the two passes.
// Each then forms Z = X R1^{-1} for its row chunks (h x b, ldz) and the
// Gram partial of Z into Gp[blockIdx.x]; workgroup 0 stores R1 (padded,
// ld PB) and the breakdown flag.
// A pivot d_i <= 1e-12 G_ii (column i nearly in the span of the previous
// ones: kappa(X) beyond ~1e6, where two CholeskyQR passes no longer give an
// orthonormal factor to working accuracy), a zero column or a non-finite
// value raises the flag; every workgroup then stops, and the rest of the
// sequence leaves the panel to the column-by-column fallback.
__global__ __launch_bounds__(small::NT) void k_apply_gram(int64_t h, int b,
                                                          double *__restrict__ G1,
                                                          double *__restrict__ X,
                                                          int64_t ldx, double *__restrict__ R1,
                                                          double *__restrict__ Z, int64_t ldz,
                                                          double *__restrict__ Gp,
                                                          int *__restrict__ flag,
                                                          double *__restrict__ tau,
                                                          double *__restrict__ T, int64_t ldt,
                                                          int *__restrict__ bar) {
    __shared__ small::Lds Xs[RC], Ms[PB];
    __shared__ double thr[PB];
    __shared__ int fail;
    __shared__ FbLds fbl;
    const int tid = threadIdx.x, q = small::wave_id();
    if (tid == 0) fail = 0;
    small::lds_load(Xs, G1, PB, b);
    __syncthreads();
    if (tid < PB) thr[tid] = tid < b ? 1e-12 * Xs[tid][tid] : 0.0;
    __syncthreads();
    // R1 -> Xs, R1^{-1} -> Ms; ends with a barrier
    sb::chol_inv(Xs, Ms, b, [&](int i) { return thr[i]; }, &fail);
    if (blockIdx.x == 0) {
        small::lds_store(Xs, R1, PB, PB, PB);
        if (tid == 0) *flag = fail;
    }
    if (fail) {  // uniform: written before chol_inv's last barrier
        // every workgroup has read G1 and (workgroup 0) stored R1 above
        if ((int64_t)gridDim.x * kFbMaxChunks * RC >= h)
            fallback_grid(h, b, X, ldx, Z, Gp, G1, tau, T, ldt, flag, bar, Xs, Ms, fbl);
        return;
    }
    sb::v4d g[3] = {sb::zero4(), sb::zero4(), sb::zero4()};
    for (int64_t r0 = (int64_t)blockIdx.x * RC; r0 < h; r0 += (int64_t)gridDim.x * RC) {
        __syncthreads();
        chunk_load(Xs, X, ldx, r0, h, b);
        __syncthreads();
        sb::v4d acc[4];
        const int64_t row = r0 + 16 * q + lrow();
        rowtile_upper(acc, Xs, Ms, q);
#pragma unroll
        for (int jb = 0; jb < 4; ++jb) {
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int col = 16 * jb + lcol(r);
                if (row < h && col < b) Z[row + (int64_t)col * ldz] = acc[jb][r];
            }
        }
        __syncthreads();
#pragma unroll
        for (int jb = 0; jb < 4; ++jb) tstoreT(Xs, q, jb, acc[jb]);
        __syncthreads();
        gram_acc(Xs, g);
    }
    gram_store(g, Gp + (size_t)blockIdx.x * PB * PB);
}

// First pass of a full panel (b = 64) fused with the previous panel's update
// of its columns, one 64-row chunk per workgroup (h <= 256 x 64).
//
// The previous panel's reflectors (Y, and W = V^T C of this panel's columns
// C, in the previous panel's frame of hc = h + 64 rows) change C to
// X = C - Y W.  They are orthogonal on those rows, so the Gram matrix of X's
// rows below the top 64 (this panel, the first CholeskyQR pass's G1) is
// Gold - Xt^T Xt, with Gold = C^T C over the frame (formed beside W by
// k_v2_formy) and Xt = X's top 64 rows (this panel's R12 block above the
// diagonal).  Every workgroup forms Xt and G1 itself (the same operations on
// the same data) and factors G1 as k_apply_gram does, so the panel's own
// Gram pass and its reduction are not needed.  The identity loses accuracy
// where most of a column lies in the previous panel's span: a pivot
// d_i <= 1e-10 Gold_ii (or 1e-12 G1_ii, k_apply_gram's test) counts as a
// breakdown.  The first pass only preconditions X; the second (k_recon,
// from Z's own Gram) sets the factor's accuracy.
//
// C's top rows are read from Ctop (their copy, ld 64), since workgroup 0
// overwrites them with Xt.
// Writes: Xt to C's top rows (TC), X to the panel X (FP64, ldx; for an FP64
// C, X = C + 64), R1, Z and the Gram partials of Z, the flag, and flag[1]:
// with one_pass_ok, 1 when Z needs no second pass (then no partials are
// written; see the test after the Cholesky), else 0.  Breakdown:
// all workgroups store X, meet at a grid barrier and take the grid
// fallback of k_apply_gram; a cancelled barrier leaves the flag at 1 for
// k_recon.
template <typename TC>
__global__ __launch_bounds__(small::NT) void k_upd_apply(int64_t h, const double *__restrict__ Y,
                                                         int64_t ldy,
                                                         const double *__restrict__ WG,
                                                         const double *__restrict__ Ctop,
                                                         TC *__restrict__ C, int64_t ldc,
                                                         double *__restrict__ X, int64_t ldx,
                                                         double *__restrict__ G1,
                                                         double *__restrict__ R1,
                                                         double *__restrict__ Z, int64_t ldz,
                                                         double *__restrict__ Gp,
                                                         int *__restrict__ flag,
                                                         double *__restrict__ tau,
                                                         double *__restrict__ T, int64_t ldt,
                                                         int *__restrict__ bar, int one_pass_ok) {
    __shared__ small::Lds As[RC], Bs[PB];
    __shared__ double thr[PB];
    __shared__ double red[3][small::NT / 64];
    __shared__ int fail;
    __shared__ FbLds fbl;
    constexpr int b = PB;
    const int tid = threadIdx.x, q = small::wave_id();
    const int l = small::lane_id(), li = l & 15, kq = l >> 4;
    const int64_t hc = h + PB, r0 = (int64_t)blockIdx.x * RC;
    const double *Gold = WG + PB * PB;
    if (tid == 0) fail = 0;
    // acc -= Y[rows of As] W: the update of one 64-row slice (row-per-lane)
    auto update = [&](sb::v4d (&acc)[4]) {
#pragma unroll
        for (int k = 0; k < 16; ++k) {
            const int kk = 4 * k + kq;
            const double av = -As[16 * q + li][kk];
#pragma unroll
            for (int jb = 0; jb < 4; ++jb)
                acc[jb] = __builtin_amdgcn_mfma_f64_16x16x4f64(Bs[kk][16 * jb + li], av, acc[jb],
                                                               0, 0, 0);
        }
    };
    // Xt = C[0:64] - Y[0:64] W, C's top rows from their copy; this
    // workgroup's chunk of Y and C is fetched meanwhile
    constexpr int TV = small::TV;
    sb::v4d xt[4], xc[4];
    double yc[TV];
    {
        double wv[TV], yv[TV];
        small::pair_fetch(wv, WG, PB, PB, PB);
        chunk_fetch(yv, Y, ldy, 0, hc, PB);
#pragma unroll
        for (int jb = 0; jb < 4; ++jb)
#pragma unroll
            for (int r = 0; r < 4; ++r)
                xt[jb][r] = Ctop[16 * q + lrow() + (16 * jb + lcol(r)) * PB];
        small::pair_stash(Bs, wv);
        small::pair_stash(As, yv);
    }
    chunk_fetch(yc, Y, ldy, PB + r0, hc, PB);
    {
        const int64_t row = PB + r0 + 16 * q + lrow();
#pragma unroll
        for (int jb = 0; jb < 4; ++jb)
#pragma unroll
            for (int r = 0; r < 4; ++r)
                xc[jb][r] = row < hc ? (double)C[row + (int64_t)(16 * jb + lcol(r)) * ldc] : 0.0;
    }
    __syncthreads();
    update(xt);
    if (blockIdx.x == 0) {
#pragma unroll
        for (int jb = 0; jb < 4; ++jb)
#pragma unroll
            for (int r = 0; r < 4; ++r)
                C[16 * q + lrow() + (int64_t)(16 * jb + lcol(r)) * ldc] = (TC)xt[jb][r];
    }
    // this workgroup's chunk: X = C - Y W on rows 64 + r0 .. of the frame
    __syncthreads();
    small::pair_stash(As, yc);
    __syncthreads();
    update(xc);
    {
        const int64_t row = r0 + 16 * q + lrow();
#pragma unroll
        for (int jb = 0; jb < 4; ++jb)
#pragma unroll
            for (int r = 0; r < 4; ++r)
                if (row < h) X[row + (int64_t)(16 * jb + lcol(r)) * ldx] = xc[jb][r];
    }
    // G1 = Gold - Xt^T Xt -> As (upper tiles; the strictly lower ones zero)
    double gtr = 0.0;  // this thread's share of trace(Gold)
    __syncthreads();
#pragma unroll
    for (int jb = 0; jb < 4; ++jb) tstoreT(As, q, jb, xt[jb]);
    __syncthreads();
    {
        // Gold's tiles (and its diagonal for the pivot thresholds) are
        // fetched before the Gram products
        double go[3][4];
#pragma unroll
        for (int u = 0; u < 3; ++u) {
            const int t = q + 4 * u;
            int i = 0, j = 0;
            if (t < 10) upper_tile(t, i, j);
#pragma unroll
            for (int r = 0; r < 4; ++r)
                go[u][r] = t < 10 ? Gold[16 * i + lrow() + (16 * j + lcol(r)) * PB] : 0.0;
        }
        const double gdiag = tid < PB ? Gold[tid * (PB + 1)] : 0.0;
        gtr = gdiag;
        sb::v4d g[3] = {sb::zero4(), sb::zero4(), sb::zero4()};
        gram_acc(As, g);
        if (tid < PB) thr[tid] = 1e-10 * gdiag;
        __syncthreads();
#pragma unroll
        for (int u = 0; u < 3; ++u) {
            const int t = q + 4 * u;
            if (t < 10) {
                int i, j;
                upper_tile(t, i, j);
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int gr = 16 * i + lrow(), gc = 16 * j + lcol(r);
                    As[gr][gc] = go[u][r] - g[u][r];
                }
            }
            if (t < 6) {
                int i, j;
                lower_tile(t, i, j);
#pragma unroll
                for (int r = 0; r < 4; ++r) As[16 * i + lrow()][16 * j + lcol(r)] = 0.0;
            }
        }
    }
    __syncthreads();
    if (tid < PB) thr[tid] = fmax(1e-12 * As[tid][tid], thr[tid]);
    __syncthreads();
    // R1 -> As, R1^{-1} -> Bs; ends with a barrier
    sb::chol_inv(As, Bs, b, [&](int i) { return thr[i]; }, &fail);
    // One pass suffices when Z = X R1^{-1} is predicted orthonormal to 1e-9:
    // G1's rounding, at most about u (tr Gold + tr G1) with the cancellation
    // of Gold - Xt^T Xt, and CholeskyQR's own u kappa^2 term are both
    // bounded by u |R1^{-1}|_F^2 (|R1|_F^2 + 2 tr Gold) (Frobenius norms
    // over-estimate the 2-norms).  Every workgroup sums the same values in
    // the same order, so all reach the same decision.
    bool one = false;
    if (one_pass_ok && !fail) {
        double sr = 0.0, si = 0.0;
        for (int e = tid; e < PB * PB; e += small::NT) {
            const int i = e / PB, j = e % PB;
            if (i <= j) {
                sr = fma(As[i][j], As[i][j], sr);
                si = fma(Bs[i][j], Bs[i][j], si);
            }
        }
        sr = small::wave_sum(sr);
        si = small::wave_sum(si);
        const double gt = small::wave_sum(gtr);
        if (l == 0) red[0][q] = sr, red[1][q] = si, red[2][q] = gt;
        __syncthreads();
        double a = 0.0, c = 0.0, t = 0.0;
#pragma unroll
        for (int w = 0; w < small::NT / 64; ++w) a += red[0][w], c += red[1][w], t += red[2][w];
        one = 0x1p-53 * c * (a + 2.0 * t) <= 1e-9;  // false for NaN
    }
    if (blockIdx.x == 0) {
        small::lds_store(As, R1, PB, PB, PB);
        if (tid == 0) *flag = fail, flag[1] = one ? 1 : 0;
    }
    if (fail) {  // uniform: written before chol_inv's last barrier
        // every workgroup's rows of X are stored before the fallback reads them
        if (!grid_sync(bar + kFbCounters - 1, (int)gridDim.x, &fbl.ok)) return;
        fallback_grid(h, b, X, ldx, Z, Gp, G1, tau, T, ldt, flag, bar, As, Bs, fbl);
        return;
    }
    // Z = X R1^{-1} for the chunk, and its Gram partial.  X comes back from
    // global memory (each lane reads what it stored; rows past h are zero)
    // rather than staying in registers through the Cholesky.
    const int64_t row = r0 + 16 * q + lrow();
#pragma unroll
    for (int jb = 0; jb < 4; ++jb)
#pragma unroll
        for (int r = 0; r < 4; ++r)
            xc[jb][r] = row < h ? X[row + (int64_t)(16 * jb + lcol(r)) * ldx] : 0.0;
    __syncthreads();  // workgroup 0's reads of R1 are done
#pragma unroll
    for (int jb = 0; jb < 4; ++jb) tstoreT(As, q, jb, xc[jb]);
    __syncthreads();
    sb::v4d acc[4];
    rowtile_upper(acc, As, Bs, q);
#pragma unroll
    for (int jb = 0; jb < 4; ++jb)
#pragma unroll
        for (int r = 0; r < 4; ++r)
            if (row < h) Z[row + (int64_t)(16 * jb + lcol(r)) * ldz] = acc[jb][r];
    if (one) return;  // uniform: no second pass, so no Gram partial of Z
    __syncthreads();
#pragma unroll
    for (int jb = 0; jb < 4; ++jb) tstoreT(As, q, jb, row < h ? acc[jb] : sb::zero4());
    __syncthreads();
    sb::v4d g[3] = {sb::zero4(), sb::zero4(), sb::zero4()};
    gram_acc(As, g);
    gram_store(g, Gp + (size_t)blockIdx.x * PB * PB);
}

// One 64-row chunk per workgroup: V(b:h) = Z(b:h) Minv into the panel A
// (unless *flag: then the fallback has written V), and, when Yf, the block
// reflector's left factor Yf (ld h) = V T^T with V = [unit-lower V1; V2].
// With Wp, also the chunk's partial of W = V^T Cn (the next block's bn
// columns at the panel's rows, ldcn) into Wp[blockIdx.x] (64 x 64, ld 64).
// With ctop, the partial holds [W | Cn^T Cn] (stride 2 x 64 x 64): the Gram
// of the next block's columns before this panel's update, from which
// k_upd_apply forms the next panel's first Gram matrix; workgroup 0 also
// copies Cn's top 64 rows to ctop (64 x 64, ld 64), so that k_upd_apply's
// workgroups read them there while its workgroup 0 overwrites them in Cn.
// zbar (when given) is zeroed by workgroup 0: the barrier counters of the
// next panel's first pass.
// Two 64 x 64 matrices of LDS (67 KB: the chunk, then Minv, T and the Cn
// chunk in turn), so the kernel fits beside one FP64 update tile on a CU.
template <typename TC>
__global__ __launch_bounds__(small::NT) void k_v2_formy(int64_t h, int b,
                                                        const double *__restrict__ Z,
                                                        const double *__restrict__ Minv,
                                                        double *__restrict__ A, int64_t lda,
                                                        const double *__restrict__ T, int64_t ldt,
                                                        double *__restrict__ Yf,
                                                        const TC *__restrict__ Cn,
                                                        int64_t ldcn, int bn,
                                                        double *__restrict__ Wp,
                                                        const int *__restrict__ flag,
                                                        float *__restrict__ V32,
                                                        float *__restrict__ Y32,
                                                        double *__restrict__ Rout, int64_t ldr,
                                                        double *__restrict__ ctop,
                                                        int *__restrict__ zbar) {
    const bool gold = ctop != nullptr;
    __shared__ small::Lds Vs[RC], Ms[PB];
    if (zbar && blockIdx.x == 0 && threadIdx.x < kFbCounters) zbar[threadIdx.x] = 0;
    const int q = small::wave_id();
    const bool fb = *flag != 0;
    const int64_t r0 = (int64_t)blockIdx.x * RC;
    const int64_t row = r0 + 16 * q + lrow();
    // loads that depend on nothing this kernel writes are issued early: T,
    // the Cn chunk and V1's rows go out while V2 is formed
    constexpr int TV = small::TV;
    double tv[TV], cn[TV], v1v[TV];
    auto fetch_rest = [&]() {
        small::pair_fetch(tv, T, ldt, b, b);  // upper part kept when stashed
        if (Wp) chunk_fetch(cn, Cn, ldcn, r0, h, bn);
        // rows above b: the unit-lower V1 (R sits above its diagonal in A)
        if (r0 < b) chunk_fetch(v1v, A, lda, r0, (int64_t)b, b);
    };
    if (!fb) {
        {
            double mv[TV], zv[TV];
            small::pair_fetch(mv, Minv, PB, PB, PB);
            chunk_fetch(zv, Z, h, r0, h, b);
            small::pair_stash(Ms, mv);
            small::pair_stash(Vs, zv);
        }
        if (Yf) fetch_rest();
        __syncthreads();
        sb::v4d acc[4];
        rowtile_upper(acc, Vs, Ms, q);
#pragma unroll
        for (int jb = 0; jb < 4; ++jb) {
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int col = 16 * jb + lcol(r);
                if (row >= b && row < h && col < b) A[row + (int64_t)col * lda] = acc[jb][r];
            }
        }
        if (!Yf) return;
        __syncthreads();
#pragma unroll
        for (int jb = 0; jb < 4; ++jb) tstoreT(Vs, q, jb, acc[jb]);
    } else {
        if (!Yf) return;
        fetch_rest();
        chunk_load(Vs, A, lda, r0, h, b);
    }
    __syncthreads();
    if (r0 < b) {
#pragma unroll
        for (int it = 0; it < TV / 2; ++it)
#pragma unroll
            for (int u = 0; u < 2; ++u) {
                const int r = small::prow() + u, c = small::pcol(it);
                const int64_t gr = r0 + r;
                if (gr < b) {
                    const double a = v1v[2 * it + u];
                    Vs[r][c] = c < b ? (gr > c ? a : (gr == c ? 1.0 : 0.0)) : 0.0;
                    if (Rout && c < b) Rout[gr + (int64_t)c * ldr] = gr <= c ? a : 0.0;
                }
            }
    }
    // T's upper part -> Ms (Minv is no longer read)
#pragma unroll
    for (int it = 0; it < TV / 2; ++it)
#pragma unroll
        for (int u = 0; u < 2; ++u) {
            const int r = small::prow() + u, c = small::pcol(it);
            Ms[r][c] = r <= c ? tv[2 * it + u] : 0.0;
        }
    __syncthreads();
    if (V32) {
#pragma unroll
        for (int jb = 0; jb < 4; ++jb)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int col = 16 * jb + lcol(r);
                if (row < h && col < b)
                    V32[row + (int64_t)col * h] = (float)Vs[16 * q + lrow()][col];
            }
    }
    // Yf = V T^T: op(T)(k, j) = T(j, k), nonzero for k >= j
    sb::v4d acc[4];
    {
        const int l = small::lane_id(), i = l & 15, kq = l >> 4;
#pragma unroll
        for (int jb = 0; jb < 4; ++jb) acc[jb] = sb::zero4();
#pragma unroll
        for (int k = 0; k < 16; ++k) {
            const int kk = 4 * k + kq;
            const double av = Vs[16 * q + i][kk];
#pragma unroll
            for (int jb = 0; jb < 4; ++jb)
                if (jb <= (k >> 2))
                    acc[jb] = __builtin_amdgcn_mfma_f64_16x16x4f64(Ms[16 * jb + i][kk], av, acc[jb],
                                                                   0, 0, 0);
        }
    }
#pragma unroll
    for (int jb = 0; jb < 4; ++jb) {
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int col = 16 * jb + lcol(r);
            if (row < h && col < b) {
                Yf[row + (int64_t)col * h] = acc[jb][r];
                if (Y32) Y32[row + (int64_t)col * h] = (float)acc[jb][r];
            }
        }
    }
    if (!Wp) return;
    __syncthreads();  // T no longer read
#pragma unroll
    for (int it = 0; it < TV / 2; ++it)
#pragma unroll
        for (int u = 0; u < 2; ++u) {
            const int r = small::prow() + u, c = small::pcol(it);
            Ms[r][c] = cn[2 * it + u];
            if (gold && blockIdx.x == 0) ctop[r + PB * c] = cn[2 * it + u];
        }
    __syncthreads();
    // W partial: tile row q of V^T Cn over the chunk's 64 rows
    {
        const int l = small::lane_id(), i = l & 15, kq = l >> 4;
#pragma unroll
        for (int jb = 0; jb < 4; ++jb) acc[jb] = sb::zero4();
#pragma unroll
        for (int k = 0; k < 16; ++k) {
            const int kk = 4 * k + kq;
            const double av = Vs[kk][16 * q + i];
#pragma unroll
            for (int jb = 0; jb < 4; ++jb)
                acc[jb] = __builtin_amdgcn_mfma_f64_16x16x4f64(Ms[kk][16 * jb + i], av, acc[jb], 0,
                                                               0, 0);
        }
    }
    double *Wb = Wp + (size_t)blockIdx.x * (gold ? 2 : 1) * PB * PB;
#pragma unroll
    for (int jb = 0; jb < 4; ++jb)
#pragma unroll
        for (int r = 0; r < 4; ++r) Wb[16 * q + lrow() + (16 * jb + lcol(r)) * PB] = acc[jb][r];
    if (gold) {
        sb::v4d g[3] = {sb::zero4(), sb::zero4(), sb::zero4()};
        gram_acc(Ms, g);
        gram_store(g, Wb + PB * PB);
    }
}

struct CqrWs {
    double *Z, *R1, *G1, *Mi, *Gp;
    int *flag;
    int *bar;  // the fallback's barrier counters
};

// The flag and the barrier counters sit at the front, at the same address
// for every panel height, so one panel's last launch can clear the next
// panel's counters.
constexpr size_t kCqrHead = 1024;  // bytes: flag (64 ints), counters (kFbCounters ints)
static_assert(sizeof(int) * (64 + kFbCounters) <= kCqrHead, "workspace head");

CqrWs carve(void *ws, int64_t h) {
    CqrWs w;
    w.flag = (int *)ws;
    w.bar = w.flag + 64;
    w.Z = (double *)((char *)ws + kCqrHead);
    w.R1 = w.Z + (size_t)h * PB;
    w.G1 = w.R1 + PB * PB;
    w.Mi = w.G1 + PB * PB;
    w.Gp = w.Mi + PB * PB;
    return w;
}

int gram_parts(int64_t rows) {
    return (int)std::min<int64_t>((rows + RC - 1) / RC, kMaxGramParts);
}

}  // namespace

// Workspace: Z (h x PB), R1, G1, Minv (PB x PB), the Gram partials, flag
// and the fallback's barrier counters.
size_t panel_cholqr_ws_bytes(int64_t h) {
    return kCqrHead +
           sizeof(double) * ((size_t)h * PB + 3 * PB * PB + (size_t)kMaxGramParts * PB * PB);
}

// First pass: C (hc x bn) -= Y W when kin > 0, and the Gram matrix G1 of
// rows [g0, hc) of C (the panel), whose Cholesky factor the second pass
// forms.
int cholqr_gram1(hipStream_t s, int64_t hc, int bn, int kin, const double *Y, int64_t ldy,
                 const double *W, int64_t ldw, void *C, int64_t ldc, int64_t g0, void *ws,
                 bool c_f32, double *Xo, int64_t ldxo) {
    const CqrWs w = carve(ws, hc - g0);
    const int parts = gram_parts(hc);
    {
        const double hh = (double)hc;
        ProfScope ps(kin > 0 ? K_LOOKAHEAD : K_PANEL,
                     2.0 * hh * bn * kin + hh * bn * bn,
                     8.0 * hh * (kin + (kin > 0 ? 2.0 : 1.0) * bn), s);
        if (c_f32)
            hipLaunchKernelGGL(k_upd_gram<float>, dim3(parts), dim3(small::NT), 0, s, hc, kin, bn,
                               Y, ldy, W, ldw, (float *)C, ldc, g0, w.Gp, Xo, ldxo, w.bar);
        else
            hipLaunchKernelGGL(k_upd_gram<double>, dim3(parts), dim3(small::NT), 0, s, hc, kin, bn,
                               Y, ldy, W, ldw, (double *)C, ldc, g0, w.Gp, Xo, ldxo, w.bar);
        MQR_LAUNCH_CHECK();
    }
    // G1 goes to its own buffer: the next launch overwrites the partials
    return reduce_partials((int64_t)PB * PB, w.Gp, parts, (int64_t)PB * PB, s, w.G1);
}

// First Cholesky, second pass, reconstruction and gated fallback of the
// h x b panel X after cholqr_gram1 (V2 follows in cholqr_finish).
int cholqr_rest_a(hipStream_t s, int64_t h, int b, double *X, int64_t ld, double *tau, double *T,
                  int64_t ldt, void *ws) {
    const CqrWs w = carve(ws, h);
    const int parts = gram_parts(h);
    const double hh = (double)h, bb = b;
    {
        ProfScope ps(K_PANEL, hh * bb * bb + hh * bb * bb + 2.0 * bb * bb * bb / 3.0,
                     16.0 * hh * bb + 8.0 * bb * bb * 2.0, s);
        hipLaunchKernelGGL(k_apply_gram, dim3(parts), dim3(small::NT), 0, s, h, b, w.G1, X, ld,
                           w.R1, w.Z, h, w.Gp, w.flag, tau, T, ldt, w.bar);
        MQR_LAUNCH_CHECK();
    }
    MQR_CHECK(reduce_partials((int64_t)PB * PB, w.Gp, parts, (int64_t)PB * PB, s));
    {
        ProfScope ps(K_PANEL_SMALL, 8.0 * bb * bb * bb / 3.0, 8.0 * bb * bb * 6.0, s);
        hipLaunchKernelGGL(k_recon, dim3(1), dim3(small::NT), 0, s, h, b, (const double *)w.Gp,
                           (int64_t)PB, (const double *)w.Z, h, (const double *)w.R1, X, ld, tau,
                           T, ldt, w.Mi, w.G1, w.flag, nullptr);
        MQR_LAUNCH_CHECK();
    }
    return MQR_SUCCESS;
}

// V2 of the panel; with Yf, Yf (h x b, ld h) = V T^T; with Wp, also
// W = V^T Cn (b x bn, ld 64) for the next block's columns Cn (ldcn) into
// Wp[0 : 64 x 64] (Wp holds ceil(h / 64) partials), and with ctop also
// Cn^T Cn into Wp[64 x 64 : 2 x 64 x 64] (twice the partials) and Cn's top
// 64 rows into ctop (ld 64).
int cholqr_finish(hipStream_t s, int64_t h, int b, double *X, int64_t ld, const double *T,
                  int64_t ldt, double *Yf, const void *Cn, int64_t ldcn, int bn, double *Wp,
                  void *ws, bool cn_f32, float *V32, float *Y32, double *Rout, int64_t ldr,
                  double *ctop) {
    const CqrWs w = carve(ws, h);
    if (!Wp) ctop = nullptr;
    const bool gold = ctop != nullptr;
    const double hh = (double)h, bb = b;
    const unsigned chunks = (unsigned)((h + RC - 1) / RC);
    if (!Yf) Wp = nullptr, V32 = nullptr, Y32 = nullptr;
    {
        ProfScope ps(K_PANEL, hh * bb * bb + (Yf ? hh * bb * bb : 0.0) + (Wp ? 2.0 * hh * bb * bn : 0.0),
                     8.0 * hh * bb * (Yf ? 3.0 : 2.0) + (Wp ? 8.0 * hh * bn : 0.0), s);
        if (cn_f32)
            hipLaunchKernelGGL(k_v2_formy<float>, dim3(chunks), dim3(small::NT), 0, s, h, b,
                               (const double *)w.Z, (const double *)w.Mi, X, ld, T, ldt, Yf,
                               (const float *)Cn, ldcn, bn, Wp, (const int *)w.flag, V32, Y32,
                               Rout, ldr, ctop, gold ? w.bar : nullptr);
        else
            hipLaunchKernelGGL(k_v2_formy<double>, dim3(chunks), dim3(small::NT), 0, s, h, b,
                               (const double *)w.Z, (const double *)w.Mi, X, ld, T, ldt, Yf,
                               (const double *)Cn, ldcn, bn, Wp, (const int *)w.flag, V32, Y32,
                               Rout, ldr, ctop, gold ? w.bar : nullptr);
        MQR_LAUNCH_CHECK();
    }
    const int64_t part = (int64_t)(gold ? 2 : 1) * PB * PB;
    if (Wp) MQR_CHECK(reduce_partials(part, Wp, (int)chunks, part, s));
    return MQR_SUCCESS;
}

// The panel of h = hc - 64 rows at X after the previous 64-column panel's
// update (Y: hc x 64, ld ldy; WG = [W | Gold] from cholqr_finish with gold)
// of its columns C (hc x 64, ldc; FP32 when c_f32, then X is the panel's
// FP64 copy, else X = C + 64): update, first pass, second Gram and
// reconstruction in k_upd_apply + reduction + k_recon; cholqr_finish
// follows as after cholqr_rest_a.  Needs h <= 256 x 64.
int cholqr_fused_a(hipStream_t s, int64_t h, const double *Y, int64_t ldy, const double *WG,
                   const double *Ctop, void *C, int64_t ldc, bool c_f32, double *X, int64_t ldx,
                   double *tau, double *T, int64_t ldt, void *ws, bool one_pass) {
    const CqrWs w = carve(ws, h);
    const int parts = (int)((h + RC - 1) / RC);
    if (h <= 0 || parts > kMaxGramParts) return MQR_ERR_INTERNAL;
    const double hh = (double)h, bb = PB;
    {
        ProfScope ps(K_PANEL, 2.0 * (hh + bb) * bb * bb + 3.0 * hh * bb * bb,
                     8.0 * (hh + bb) * bb * (c_f32 ? 2.5 : 3.0) + 16.0 * hh * bb, s);
        if (c_f32)
            hipLaunchKernelGGL(k_upd_apply<float>, dim3(parts), dim3(small::NT), 0, s, h, Y, ldy,
                               WG, Ctop, (float *)C, ldc, X, ldx, w.G1, w.R1, w.Z, h, w.Gp, w.flag, tau,
                               T, ldt, w.bar, (int)one_pass);
        else
            hipLaunchKernelGGL(k_upd_apply<double>, dim3(parts), dim3(small::NT), 0, s, h, Y, ldy,
                               WG, Ctop, (double *)C, ldc, X, ldx, w.G1, w.R1, w.Z, h, w.Gp, w.flag, tau,
                               T, ldt, w.bar, (int)one_pass);
        MQR_LAUNCH_CHECK();
    }
    // flag[1]: k_upd_apply's decision that the second pass is not needed
    const int *skip = one_pass ? w.flag + 1 : nullptr;
    MQR_CHECK(reduce_partials((int64_t)PB * PB, w.Gp, parts, (int64_t)PB * PB, s, nullptr, skip));
    {
        ProfScope ps(K_PANEL_SMALL, 8.0 * bb * bb * bb / 3.0, 8.0 * bb * bb * 6.0, s);
        hipLaunchKernelGGL(k_recon, dim3(1), dim3(small::NT), 0, s, h, (int)PB,
                           (const double *)w.Gp, (int64_t)PB, (const double *)w.Z, h,
                           (const double *)w.R1, X, ldx, tau, T, ldt, w.Mi, w.G1, w.flag, skip);
        MQR_LAUNCH_CHECK();
    }
    return MQR_SUCCESS;
}

int cholqr_rest(hipStream_t s, int64_t h, int b, double *X, int64_t ld, double *tau, double *T,
                int64_t ldt, double *Yf, void *ws) {
    MQR_CHECK(cholqr_rest_a(s, h, b, X, ld, tau, T, ldt, ws));
    return cholqr_finish(s, h, b, X, ld, T, ldt, Yf, nullptr, 0, 0, nullptr, ws);
}

// Factor the h x b panel X (ld) in place.
int panel_cholqr(mqr_handle_t hd, int64_t h, int b, double *X, int64_t ld, double *tau,
                 double *T, int64_t ldt, void *ws) {
    MQR_CHECK(cholqr_gram1(hd->stream, h, b, 0, nullptr, 0, nullptr, 0, X, ld, 0, ws));
    return cholqr_rest(hd->stream, h, b, X, ld, tau, T, ldt, nullptr, ws);
}

}  // namespace mqr
