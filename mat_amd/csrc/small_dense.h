// Dense kernels for the b x b factors of a panel (b <= 64): Cholesky, LU with
// sign choice, triangular inverse and products of one FP64 matrix.
//
// The sequential factorisations run in a single wave with the matrix held in
// registers: lane c owns column c (64 registers, loops fully unrolled so that
// every register index is a compile-time constant).  A step needs values from
// other lanes: rows are broadcast through a 64-double LDS buffer (Cholesky)
// or read from a read-only LDS copy (inverse); single values come from
// v_readlane (LU).  Products of two matrices use all four waves through LDS.
#pragma once

#include "mqr_internal.h"

namespace mqr {
namespace small {

constexpr int N = 64;  // matrix order (smaller problems are padded with I)

__device__ __forceinline__ int lane_id() { return threadIdx.x & 63; }
__device__ __forceinline__ int wave_id() { return threadIdx.x >> 6; }

// value of v in lane `src` (a compile-time constant after unrolling)
__device__ __forceinline__ double bcast(double v, int src) {
    const int lo = __builtin_amdgcn_readlane(__double2loint(v), src);
    const int hi = __builtin_amdgcn_readlane(__double2hiint(v), src);
    return __hiloint2double(hi, lo);
}

// Column c of a column-major matrix (ld) into a[] of lane c; identity padding
// outside the leading b x b block.
__device__ __forceinline__ void load_col(double (&a)[N], const double *M, int64_t ld, int b) {
    const int c = lane_id();
#pragma unroll
    for (int r = 0; r < N; ++r) a[r] = (r < b && c < b) ? M[r + (int64_t)c * ld] : (r == c ? 1.0 : 0.0);
}

__device__ __forceinline__ void store_col(const double (&a)[N], double *M, int64_t ld) {
    const int c = lane_id();
#pragma unroll
    for (int r = 0; r < N; ++r) M[r + (int64_t)c * ld] = a[r];
}

// LDS matrix S[r][c] (row stride N + 1) <-> columns in lanes
__device__ __forceinline__ void col_to_lds(const double (&a)[N], double (*S)[N + 1]) {
    const int c = lane_id();
#pragma unroll
    for (int r = 0; r < N; ++r) S[r][c] = a[r];
}
__device__ __forceinline__ void lds_to_col(double (&a)[N], double (*S)[N + 1]) {
    const int c = lane_id();
#pragma unroll
    for (int r = 0; r < N; ++r) a[r] = S[r][c];
}

// Upper Cholesky factor in place, row i of R broadcast through LDS (rowbuf,
// N doubles).  Lanes below the diagonal are updated too (their values are
// never read and are zeroed at the end), which keeps the update unmasked:
// one broadcast load and one FMA per element.  Every thread of the block
// calls this (the barriers are uniform); only `active` threads hold data.
// *fail is set when a pivot among the first b is <= thr or NaN.
__device__ __forceinline__ void chol_lds(double (&a)[N], double *rowbuf, int b, double thr,
                                         int *fail, bool active) {
    const int c = lane_id();
#pragma unroll
    for (int i = 0; i < N; ++i) {
        double rc = 0.0;
        if (active) {
            const double d = bcast(a[i], i);
            if (c == 0 && i < b && !(d > thr)) *fail = 1;
            const double rii = sqrt(fmax(d, 0.0));
            rc = (c == i) ? rii : a[i] * (1.0 / rii);  // R(i, c) for c >= i
            rowbuf[c] = rc;
        }
        __syncthreads();
        if (active) {
#pragma unroll
            for (int r = i + 1; r < N; ++r) a[r] = fma(-rowbuf[r], rc, a[r]);
            a[i] = rc;
        }
        __syncthreads();
    }
    if (active) {
#pragma unroll
        for (int r = 0; r < N; ++r) a[r] = (r <= c) ? a[r] : 0.0;
    }
}

// x = R^{-1} for upper R given row-major in LDS (S[r][c], read only) with
// rdiag[i] = 1 / R(i, i): lane c computes column c bottom-up; four partial
// sums per row, no barriers.
__device__ __forceinline__ void inv_upper_lds(double (*S)[N + 1], const double *rdiag,
                                              double (&x)[N]) {
    const int c = lane_id();
#pragma unroll
    for (int i = N - 1; i >= 0; --i) {
        double s[4] = {(c == i) ? 1.0 : 0.0, 0.0, 0.0, 0.0};
#pragma unroll
        for (int k = i + 1; k < N; ++k) s[k & 3] = fma(-S[i][k], x[k], s[k & 3]);
        x[i] = ((s[0] + s[1]) + (s[2] + s[3])) * rdiag[i];
    }
}

// LU without pivoting of M - S with S_j = -sign(current diagonal), sign(0) = +:
// on exit a holds L (unit, strictly below) and U (on and above); sg[j] (LDS,
// written by lane 0) the signs.
__device__ __forceinline__ void lu_sign(double (&a)[N], double *sg) {
    const int c = lane_id();
#pragma unroll
    for (int j = 0; j < N; ++j) {
        const double aj = bcast(a[j], j);
        const double sj = aj >= 0.0 ? -1.0 : 1.0;
        const double piv = aj - sj;
        const double ipiv = 1.0 / piv;  // |piv| >= 1 by the sign choice
        if (c == 0) sg[j] = sj;
#pragma unroll
        for (int r = j + 1; r < N; ++r) {
            const double l = bcast(a[r], j) * ipiv;  // L(r, j)
            a[r] = (c > j) ? fma(-l, a[j], a[r]) : ((c == j) ? l : a[r]);
        }
        a[j] = (c == j) ? piv : a[j];
    }
}

// C = A * B with A, B in LDS (S[r][c]); all four waves compute, wave w rows
// 16 w .. 16 w + 15 of column lane c.  upper: both triangular, k in [r, c].
// C may alias A (each row of C is written by the wave that read that row).
__device__ __forceinline__ void matmul_lds(double (*SA)[N + 1], double (*SB)[N + 1],
                                           double (*SC)[N + 1], bool upper) {
    const int c = lane_id(), w = wave_id();
#pragma unroll 1
    for (int q = 0; q < 16; ++q) {
        const int r = w * 16 + q;
        double acc = 0.0;
        const int k0 = upper ? r : 0, k1 = upper ? c : N - 1;
        for (int k = k0; k <= k1; ++k) acc = fma(SA[r][k], SB[k][c], acc);
        SC[r][c] = acc;
    }
}

}  // namespace small
}  // namespace mqr
