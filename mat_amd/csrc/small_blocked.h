// Blocked dense kernels for the b x b factors of a panel (b <= 64) in one
// workgroup of 256 threads (four waves): Cholesky with the inverse of its
// factor, LU with sign choice with the inverses of both factors, and
// products.
//
// The 64 x 64 matrices live in LDS (row r at S[r]) and are processed in
// 16 x 16 tiles.  A factorisation goes by block columns: one wave factors
// the 16 x 16 diagonal block in registers (lane j holds column j; each
// elimination step broadcasts within 16-lane rows by DPP, no barrier) and
// inverts its triangular factors by substitution, one column per lane; then
// the four waves update the block row/column and the trailing tiles on the
// FP64 MFMA.
// The full triangular inverses are assembled from the diagonal-block
// inverses by recursive doubling, X_12 = -X_11 R_12 X_22, in two levels of
// tile products.  Every barrier is reached by all threads.
#pragma once

#include "small_dense.h"

namespace mqr {
namespace sb {

using small::Lds;
constexpr int N = small::N;
typedef double v4d __attribute__((ext_vector_type(4)));

__device__ __forceinline__ v4d zero4() { return (v4d){0.0, 0.0, 0.0, 0.0}; }

// acc += s * op(A)[tile ia, ka] * op(B)[tile kb, jb]; op(X) = X^T when the
// flag is set (tile (i, k) of X^T is the transpose of tile (k, i) of X).
// Lane l, register r of the result: row 16 ia + (l >> 4) + 4 r, column
// 16 jb + (l & 15).
template <bool TA, bool TB>
__device__ __forceinline__ v4d tmma(v4d acc, const Lds *A, int ia, int ka, const Lds *B, int kb,
                                    int jb, double s = 1.0) {
    const int l = small::lane_id(), i = l & 15, kq = l >> 4;
#pragma unroll
    for (int k4 = 0; k4 < 4; ++k4) {
        const int ak = 16 * ka + 4 * k4 + kq, bk = 16 * kb + 4 * k4 + kq;
        const double av = TA ? A[ak][16 * ia + i] : A[16 * ia + i][ak];
        const double bv = TB ? B[16 * jb + i][bk] : B[bk][16 * jb + i];
        acc = __builtin_amdgcn_mfma_f64_16x16x4f64(s * av, bv, acc, 0, 0, 0);
    }
    return acc;
}

// acc += A[tile ia, k] * B[tile k, jb] with A's tile held by the calling
// lane as the MFMA operand: a[k4] = A[16 ia + (l & 15)][16 k + 4 k4 + (l >> 4)]
// (the values tmma<false, ...> reads from LDS).
__device__ __forceinline__ v4d tmma_reg(v4d acc, const double (&a)[4], const Lds *B, int kb,
                                        int jb) {
    const int l = small::lane_id(), i = l & 15, kq = l >> 4;
#pragma unroll
    for (int k4 = 0; k4 < 4; ++k4)
        acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a[k4], B[16 * kb + 4 * k4 + kq][16 * jb + i], acc,
                                                   0, 0, 0);
    return acc;
}

__device__ __forceinline__ int trow(int r) { return (small::lane_id() >> 4) + 4 * r; }
__device__ __forceinline__ int tcol() { return small::lane_id() & 15; }

__device__ __forceinline__ v4d tload(const Lds *C, int ib, int jb) {
    v4d v;
#pragma unroll
    for (int r = 0; r < 4; ++r) v[r] = C[16 * ib + trow(r)][16 * jb + tcol()];
    return v;
}

__device__ __forceinline__ void tstore(Lds *C, int ib, int jb, v4d v) {
#pragma unroll
    for (int r = 0; r < 4; ++r) C[16 * ib + trow(r)][16 * jb + tcol()] = v[r];
}

__device__ __forceinline__ void tzero(Lds *C, int ib, int jb) { tstore(C, ib, jb, zero4()); }

// 1 / sqrt(d): the hardware estimate y0 (relative error below 6e-8 on
// gfx950) corrected to second order, y0 (1 - e)^{-1/2} = y0 (1 + e / 2 +
// 3 e^2 / 8 + ...) with e = 1 - d y0^2; the dropped terms are below 2^-70
// and the result is within an ulp.  Four dependent operations after the
// estimate (two Newton steps take six).
__device__ __forceinline__ double rsqrt_nr(double d) {
    const double y0 = __builtin_amdgcn_rsq(d);
    const double e = fma(-d * y0, y0, 1.0);
    return fma(y0 * e, fma(0.375, e, 0.5), y0);
}

// Value of v in lane K of the calling lane's 16-lane row (DPP row_newbcast).
// The diagonal-block routines keep column l & 15 in lane l, so every row of
// the wave holds the same 16 columns and this is a broadcast from column K.
template <int K>
__device__ __forceinline__ double rbcast(double v) {
    const int lo = __builtin_amdgcn_mov_dpp(__double2loint(v), 0x150 + K, 0xF, 0xF, false);
    const int hi = __builtin_amdgcn_mov_dpp(__double2hiint(v), 0x150 + K, 0xF, 0xF, false);
    return __hiloint2double(hi, lo);
}

// The sixteen elimination steps of diag_chol, software-pipelined: step K
// first updates column K + 1 (the next pivot's column), starts the next
// pivot's reciprocal square root, and spreads the rest of its rank-one
// update between that root's correction steps, so that the in-order issue
// fills the pivot chain's latency.  On entry to step K, yk = 1 / sqrt(pivot K).
template <int I0, int I1, int K>
__device__ __forceinline__ void chol_upd_range(double (&a)[16]) {
    if constexpr (I0 < I1 && I0 < 16) {
        a[I0] = fma(-rbcast<I0 < 16 ? I0 : 15>(a[K]), a[K], a[I0]);
        chol_upd_range<I0 + 1, I1, K>(a);
    }
}
template <int K, typename Thr>
__device__ __forceinline__ void chol_steps_pipe(double (&a)[16], double (&y)[16], int o, int b,
                                                Thr thr, bool &bad, double yk) {
    if constexpr (K < 16) {
        y[K] = yk;
        a[K] *= yk;  // row K of R (columns j >= K)
        if constexpr (K + 1 < 16) {
            a[K + 1] = fma(-rbcast<K + 1>(a[K]), a[K], a[K + 1]);
            const double dn = rbcast<K + 1>(a[K + 1]);
            // rsqrt_nr(dn), its operations interleaved with the updates
            const double r0 = __builtin_amdgcn_rsq(dn);
            chol_upd_range<K + 2, K + 5, K>(a);
            const double e = fma(-dn * r0, r0, 1.0);
            chol_upd_range<K + 5, K + 8, K>(a);
            const double r = fma(r0 * e, fma(0.375, e, 0.5), r0);
            chol_upd_range<K + 8, 16, K>(a);
            bad |= (o + K + 1 < b) && !(dn > thr(o + K + 1));
            chol_steps_pipe<K + 1>(a, y, o, b, thr, bad, r);
        }
    }
}

// Row updates I0 <= i < I1 of step K of diag_lu (lik = L(i, K) from lane K)
template <int I0, int I1, int K>
__device__ __forceinline__ void lu_upd_range(double (&a)[16], int j, double ipk) {
    if constexpr (I0 < I1 && I0 < 16) {
        const double lik = rbcast<K>(a[I0]) * ipk;
        a[I0] = (j > K) ? fma(-lik, a[K], a[I0]) : ((j == K) ? lik : a[I0]);
        lu_upd_range<I0 + 1, I1, K>(a, j, ipk);
    }
}

// The sixteen steps of diag_lu, software-pipelined like chol_steps_pipe:
// step K updates row K + 1 first, starts the next pivot's reciprocal and
// spreads its other row updates between that reciprocal's correction
// steps.  On entry piv is pivot K (after its sign shift), s its sign and
// ipk = 1 / piv.
template <int K>
__device__ __forceinline__ void lu_steps_pipe(double (&a)[16], double (&ip)[16], int j, int o,
                                              double *sg, double piv, double s, double ipk) {
    if constexpr (K < 16) {
        ip[K] = ipk;
        if (small::lane_id() == 0) sg[o + K] = s;
        if (j == K) a[K] = piv;
        if constexpr (K + 1 < 16) {
            lu_upd_range<K + 1, K + 2, K>(a, j, ipk);
            const double akk = rbcast<K + 1>(a[K + 1]);
            const double sn = akk >= 0.0 ? -1.0 : 1.0;
            const double pn = akk - sn;
            // small::recip(pn), its operations interleaved with the updates
            const double y0 = __builtin_amdgcn_rcp(pn);
            lu_upd_range<K + 2, K + 5, K>(a, j, ipk);
            const double e = fma(-pn, y0, 1.0);
            lu_upd_range<K + 5, K + 8, K>(a, j, ipk);
            const double y = fma(y0 * e, 1.0 + e, y0);
            lu_upd_range<K + 8, 16, K>(a, j, ipk);
            lu_steps_pipe<K + 1>(a, ip, j, o, sg, pn, sn, y);
        }
    }
}

// One wave: upper Cholesky factor of the 16 x 16 diagonal block kb of S
// (upper triangle read, Schur-updated by the earlier blocks) in place, zero
// below its diagonal; its inverse into the same block of X.  A pivot of
// row i < b (the Schur diagonal before the square root) <= thr(i), or NaN,
// sets *fail.
template <typename Thr>
__device__ __forceinline__ void diag_chol(Lds *S, Lds *X, int kb, int b, Thr thr, int *fail) {
    const int j = small::lane_id() & 15, o = 16 * kb;
    double a[16], y[16], x[16];
#pragma unroll
    for (int i = 0; i < 16; ++i) a[i] = S[o + i][o + j];
    bool bad = false;
    {
        const double d = rbcast<0>(a[0]);
        bad |= (o < b) && !(d > thr(o));
        chol_steps_pipe<0>(a, y, o, b, thr, bad, rsqrt_nr(d));
    }
    if (bad && small::lane_id() == 0) *fail = 1;
    if (small::lane_id() < 16) {
#pragma unroll
        for (int i = 0; i < 16; ++i) S[o + i][o + j] = (i <= j) ? a[i] : 0.0;
    }
    // the wave's own LDS writes are seen by its later reads
    __builtin_amdgcn_wave_barrier();
    // column j of R^{-1}: back substitution on e_j (R read back as LDS
    // broadcasts)
#pragma unroll
    for (int i = 0; i < 16; ++i) x[i] = (i == j) ? 1.0 : 0.0;
#pragma unroll
    for (int k = 15; k >= 0; --k) {
        x[k] *= y[k];
#pragma unroll
        for (int i = 0; i < k; ++i) x[i] = fma(-S[o + i][o + k], x[k], x[i]);
    }
    if (small::lane_id() < 16) {
#pragma unroll
        for (int i = 0; i < 16; ++i) X[o + i][o + j] = (i <= j) ? x[i] : 0.0;
    }
}

// One wave: LU without pivoting of the diagonal block kb of S - diag(sg),
// sg chosen step by step as -sign of the current diagonal (sign(0) = +), so
// |U_ii| >= 1.  L (unit, strictly below) and U in place; ip[i] = 1 / U_ii
// stays in the wave's registers for diag_inv_u.
__device__ __forceinline__ void diag_lu(Lds *S, int kb, double *sg, double (&ip)[16]) {
    const int j = small::lane_id() & 15, o = 16 * kb;
    double a[16];
#pragma unroll
    for (int i = 0; i < 16; ++i) a[i] = S[o + i][o + j];
    {
        const double akk = rbcast<0>(a[0]);
        const double s = akk >= 0.0 ? -1.0 : 1.0;
        const double piv = akk - s;
        lu_steps_pipe<0>(a, ip, j, o, sg, piv, s, small::recip(piv));
    }
    if (small::lane_id() < 16) {
#pragma unroll
        for (int i = 0; i < 16; ++i) S[o + i][o + j] = a[i];
    }
}

// One wave: U^{-1} of the diagonal block kb of S (after diag_lu, ip from
// it) into the block of XU, column j by back substitution in lane j (U read
// back as LDS broadcasts).
__device__ __forceinline__ void diag_inv_u(const Lds *S, Lds *XU, int kb, const double (&ip)[16]) {
    const int j = small::lane_id() & 15, o = 16 * kb;
    double x[16];
#pragma unroll
    for (int i = 0; i < 16; ++i) x[i] = (i == j) ? 1.0 : 0.0;
#pragma unroll
    for (int k = 15; k >= 0; --k) {
        x[k] *= ip[k];
#pragma unroll
        for (int i = 0; i < k; ++i) x[i] = fma(-S[o + i][o + k], x[k], x[i]);
    }
    if (small::lane_id() < 16) {
#pragma unroll
        for (int i = 0; i < 16; ++i) XU[o + i][o + j] = (i <= j) ? x[i] : 0.0;
    }
}

// One wave: L^{-1} (unit lower, ones on the diagonal) of the diagonal block
// kb of S (after diag_lu) into tile (kb, jt) of XL, column j by forward
// substitution in lane j.
__device__ __forceinline__ void diag_inv_l(const Lds *S, Lds *XL, int kb, int jt) {
    const int j = small::lane_id() & 15, o = 16 * kb, ot = 16 * jt;
    double x[16];
#pragma unroll
    for (int i = 0; i < 16; ++i) x[i] = (i == j) ? 1.0 : 0.0;
#pragma unroll
    for (int k = 0; k < 16; ++k) {
#pragma unroll
        for (int i = k + 1; i < 16; ++i) x[i] = fma(-S[o + i][o + k], x[k], x[i]);
    }
    if (small::lane_id() < 16) {
#pragma unroll
        for (int i = 0; i < 16; ++i) XL[o + i][ot + j] = (i >= j) ? x[i] : 0.0;
    }
}

// X = R^{-1} for the upper-triangular R whose strictly upper tiles are read
// from R, given the inverses of its diagonal tiles in X's diagonal tiles.
// X's strictly lower tiles serve as scratch and are zero on exit.
__device__ __forceinline__ void inv_upper_dbl(const Lds *R, Lds *X) {
    const int q = small::wave_id();
    // level 1: X[2p][2p+1] = -X[2p][2p] (R[2p][2p+1] X[2p+1][2p+1])
    if (q < 2) {
        const int p = 2 * q;
        tstore(X, p + 1, p, tmma<false, false>(zero4(), R, p, p + 1, X, p + 1, p + 1));
    }
    __syncthreads();
    if (q < 2) {
        const int p = 2 * q;
        tstore(X, p, p + 1, tmma<false, false>(zero4(), X, p, p, X, p + 1, p, -1.0));
    }
    __syncthreads();
    // level 2: W[i][j] = sum_k R[i][2+k] X[2+k][2+j] (k <= j) into X[2+i][j],
    // then X[i][2+j] = -sum_{k >= i} X[i][k] W[k][j]
    const int i = q >> 1, jj = q & 1;
    {
        v4d acc = zero4();
        for (int k = 0; k <= jj; ++k) acc = tmma<false, false>(acc, R, i, 2 + k, X, 2 + k, 2 + jj);
        tstore(X, 2 + i, jj, acc);
    }
    __syncthreads();
    {
        v4d acc = zero4();
        for (int k = i; k < 2; ++k) acc = tmma<false, false>(acc, X, i, k, X, 2 + k, jj, -1.0);
        tstore(X, i, 2 + jj, acc);
    }
    __syncthreads();
    // strictly lower tiles back to zero: (1,0), (2,0), (2,1), (3,0), (3,1), (3,2)
    {
        const int t0 = q, t1 = q + 4;
        const int ti[6] = {1, 2, 2, 3, 3, 3}, tj[6] = {0, 0, 1, 0, 1, 2};
        tzero(X, ti[t0], tj[t0]);
        if (t1 < 6) tzero(X, ti[t1], tj[t1]);
    }
}

// X = L^{-1} for the unit lower-triangular L whose strictly lower tiles are
// read from L, given the inverses of its diagonal tiles in X's diagonal
// tiles.  X's strictly upper tiles serve as scratch and are zero on exit.
__device__ __forceinline__ void inv_lower_dbl(const Lds *L, Lds *X) {
    const int q = small::wave_id();
    // level 1: X[2p+1][2p] = -X[2p+1][2p+1] (L[2p+1][2p] X[2p][2p])
    if (q < 2) {
        const int p = 2 * q;
        tstore(X, p, p + 1, tmma<false, false>(zero4(), L, p + 1, p, X, p, p));
    }
    __syncthreads();
    if (q < 2) {
        const int p = 2 * q;
        tstore(X, p + 1, p, tmma<false, false>(zero4(), X, p + 1, p + 1, X, p, p + 1, -1.0));
    }
    __syncthreads();
    // level 2: W[k][j] = sum_{m >= j} L[2+k][m] X[m][j] into X[j][2+k],
    // then X[2+i][j] = -sum_{k <= i} X[2+i][2+k] W[k][j]
    const int kk = q >> 1, jj = q & 1;
    {
        v4d acc = zero4();
        for (int m = jj; m < 2; ++m) acc = tmma<false, false>(acc, L, 2 + kk, m, X, m, jj);
        tstore(X, jj, 2 + kk, acc);
    }
    __syncthreads();
    {
        const int i = kk;
        v4d acc = zero4();
        // W[k][jj] sits in tile (jj, 2 + k) of X
        for (int k = 0; k <= i; ++k) acc = tmma<false, false>(acc, X, 2 + i, 2 + k, X, jj, 2 + k, -1.0);
        tstore(X, 2 + i, jj, acc);
    }
    __syncthreads();
    {
        const int ti[6] = {0, 0, 1, 0, 1, 2}, tj[6] = {1, 2, 2, 3, 3, 3};
        tzero(X, ti[q], tj[q]);
        if (q + 4 < 6) tzero(X, ti[q + 4], tj[q + 4]);
    }
}

// Upper Cholesky factor of the symmetric S (upper triangle read) in place,
// R^T R = S, zero below the diagonal; X = R^{-1}.  A pivot of row i < b
// (the Schur diagonal before its square root) <= thr(i), or NaN, sets
// *fail (LDS).
template <typename Thr>
__device__ __forceinline__ void chol_inv(Lds *S, Lds *X, int b, Thr thr, int *fail) {
    const int q = small::wave_id();
#pragma unroll 1
    for (int kb = 0; kb < 4; ++kb) {
        if (q == 0) diag_chol(S, X, kb, b, thr, fail);
        __syncthreads();
        // block row: R[kb][jb] = R_kk^{-T} S[kb][jb]
        if (q < 3 - kb) {
            const int jb = kb + 1 + q;
            tstore(S, kb, jb, tmma<true, false>(zero4(), X, kb, kb, S, kb, jb));
        }
        __syncthreads();
        // trailing: S[ib][jb] -= R[kb][ib]^T R[kb][jb], kb < ib <= jb
        int t = 0;
        for (int ib = kb + 1; ib < 4; ++ib)
            for (int jb = ib; jb < 4; ++jb, ++t)
                if ((t & 3) == q)
                    tstore(S, ib, jb, tmma<true, false>(tload(S, ib, jb), S, ib, kb, S, kb, jb, -1.0));
        __syncthreads();
    }
    // S's strictly lower tiles (still the input's lower triangle) to zero
    {
        const int ti[6] = {1, 2, 2, 3, 3, 3}, tj[6] = {0, 0, 1, 0, 1, 2};
        tzero(S, ti[q], tj[q]);
        if (q + 4 < 6) tzero(S, ti[q + 4], tj[q + 4]);
    }
    inv_upper_dbl(S, X);
    __syncthreads();
}

// LU without pivoting of S - diag(sg), the signs chosen during the
// elimination (sg[i] = -sign of the current diagonal, sign(0) = +): L (unit,
// strictly below) and U in place; X = U^{-1}.  Two matrices of LDS: the
// inverse of L's diagonal block kb, used only in its own block step, is
// kept in X's tile (kb, kb + 1) until the doubling overwrites it (block 3
// needs none).  L^{-1} follows from diag_inv_l on every diagonal block and
// inv_lower_dbl.
__device__ __forceinline__ void lu_sign_invu(Lds *S, Lds *X, double *sg) {
    const int q = small::wave_id();
#pragma unroll 1
    for (int kb = 0; kb < 4; ++kb) {
        // wave 0 factors the diagonal block; then waves 0 and 1 invert its
        // two factors side by side
        double ip[16];
        if (q == 0) diag_lu(S, kb, sg, ip);
        __syncthreads();
        if (q == 0) diag_inv_u(S, X, kb, ip);
        else if (q == 1 && kb < 3) diag_inv_l(S, X, kb, kb + 1);
        __syncthreads();
        // L tiles S[ib][kb] <- S[ib][kb] U_kk^{-1}; U tiles S[kb][jb] <- L_kk^{-1} S[kb][jb]
        const int nt = 3 - kb;
        for (int t = q; t < 2 * nt; t += 4) {
            if (t < nt) {
                const int ib = kb + 1 + t;
                tstore(S, ib, kb, tmma<false, false>(zero4(), S, ib, kb, X, kb, kb));
            } else {
                const int jb = kb + 1 + (t - nt);
                tstore(S, kb, jb, tmma<false, false>(zero4(), X, kb, kb + 1, S, kb, jb));
            }
        }
        __syncthreads();
        // trailing: S[ib][jb] -= L[ib][kb] U[kb][jb]
        for (int t = q; t < nt * nt; t += 4) {
            const int ib = kb + 1 + t / nt, jb = kb + 1 + t % nt;
            tstore(S, ib, jb, tmma<false, false>(tload(S, ib, jb), S, ib, kb, S, kb, jb, -1.0));
        }
        __syncthreads();
    }
    inv_upper_dbl(S, X);
    __syncthreads();
}

// X = L^{-1} for the unit lower L (strictly below the diagonal of S, after
// lu_sign_invu); X's previous contents are overwritten.
__device__ __forceinline__ void inv_unit_lower(const Lds *S, Lds *X) {
    const int q = small::wave_id();
    diag_inv_l(S, X, q, q);
    __syncthreads();
    inv_lower_dbl(S, X);
    __syncthreads();
}

}  // namespace sb
}  // namespace mqr
