// Dense kernels for the b x b factors of a panel (b <= 64) in one workgroup
// of 256 threads: Cholesky, LU with sign choice, triangular inverse and
// substitution, and products.
//
// Distributed layout: thread t holds column c = t & 63 of the rows
// r = 4 k + q, k = 0..15, q = t >> 6 (its wave), in a[k].  Sixteen registers
// per matrix, so several matrices stay resident without spilling.  An
// elimination step i is done by the wave that owns row i (q = i & 3, a
// compile-time register index k = i >> 2 once the loop is unrolled): it
// finishes row i and publishes it through a 64-double LDS buffer; after one
// barrier every wave updates its own rows.  The buffer alternates between
// two slots, so one barrier per step suffices: a slot is rewritten two steps
// later, after every wave has passed the barrier that follows its last read.
// All barriers are reached by every thread of the block.
//
// Products of two LDS-resident matrices run on the FP64 MFMA
// (v_mfma_f64_16x16x4f64): wave q computes output rows 16 q .. 16 q + 15.
#pragma once

#include "mqr_internal.h"

namespace mqr {
namespace small {

constexpr int N = 64;    // matrix order (smaller problems are padded with I)
constexpr int NT = 256;  // threads of the workgroup
constexpr int RK = 16;   // rows held per thread
using Lds = double[N + 1];  // row of an LDS matrix S[r][c]

typedef double v4d __attribute__((ext_vector_type(4)));

__device__ __forceinline__ int lane_id() { return threadIdx.x & 63; }
__device__ __forceinline__ int wave_id() { return threadIdx.x >> 6; }

// value of v in lane `src` of the calling wave
__device__ __forceinline__ double bcast(double v, int src) {
    const int lo = __builtin_amdgcn_readlane(__double2loint(v), src);
    const int hi = __builtin_amdgcn_readlane(__double2hiint(v), src);
    return __hiloint2double(hi, lo);
}

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// S = M(0:b, 0:b) (column-major, ld), identity outside; coalesced reads.
__device__ __forceinline__ void lds_load(Lds *S, const double *__restrict__ M, int64_t ld, int b) {
    for (int e = threadIdx.x; e < N * N; e += NT) {
        const int r = e & (N - 1), c = e >> 6;
        S[r][c] = (r < b && c < b) ? M[r + (int64_t)c * ld] : (r == c ? 1.0 : 0.0);
    }
}

// M(0:rows, 0:cols) = S; coalesced writes.
__device__ __forceinline__ void lds_store(Lds *S, double *__restrict__ M, int64_t ld, int rows,
                                          int cols) {
    for (int e = threadIdx.x; e < N * N; e += NT) {
        const int r = e & (N - 1), c = e >> 6;
        if (r < rows && c < cols) M[r + (int64_t)c * ld] = S[r][c];
    }
}

__device__ __forceinline__ void from_lds(double (&a)[RK], Lds *S) {
    const int c = lane_id(), q = wave_id();
#pragma unroll
    for (int k = 0; k < RK; ++k) a[k] = S[4 * k + q][c];
}

__device__ __forceinline__ void to_lds(const double (&a)[RK], Lds *S) {
    const int c = lane_id(), q = wave_id();
#pragma unroll
    for (int k = 0; k < RK; ++k) S[4 * k + q][c] = a[k];
}

__device__ __forceinline__ void set_identity(double (&a)[RK]) {
    const int c = lane_id(), q = wave_id();
#pragma unroll
    for (int k = 0; k < RK; ++k) a[k] = (4 * k + q == c) ? 1.0 : 0.0;
}

__device__ __forceinline__ void keep_upper(double (&a)[RK]) {
    const int c = lane_id(), q = wave_id();
#pragma unroll
    for (int k = 0; k < RK; ++k) a[k] = (4 * k + q <= c) ? a[k] : 0.0;
}

// Upper Cholesky factor R (G = R^T R) in place.  Rows below the diagonal
// are updated as well (their values are never read) and zeroed at the end.
// *fail (LDS) is set when one of the first b pivots is <= thr or NaN.
__device__ __forceinline__ void chol(double (&a)[RK], double (*rb)[N], int b, double thr,
                                     int *fail) {
    const int c = lane_id(), q = wave_id();
#pragma unroll
    for (int i = 0; i < N; ++i) {
        double *buf = rb[i & 1];
        if (q == (i & 3)) {
            const double d = bcast(a[i >> 2], i);
            if (c == 0 && i < b && !(d > thr)) *fail = 1;
            const double rii = sqrt(fmax(d, 0.0));
            const double rc = (c == i) ? rii : a[i >> 2] * (1.0 / rii);
            a[i >> 2] = rc;
            buf[c] = rc;
        }
        __syncthreads();
        const double rcv = buf[c];
#pragma unroll
        for (int k = 0; k < RK; ++k)
            if (4 * k + q > i) a[k] = fma(-buf[4 * k + q], rcv, a[k]);
    }
    keep_upper(a);
}

// x = R^{-1} for the upper-triangular R held in S (read only).
__device__ __forceinline__ void inv_upper(Lds *S, double (*rb)[N], double (&x)[RK]) {
    const int c = lane_id(), q = wave_id();
    set_identity(x);
#pragma unroll
    for (int i = N - 1; i >= 0; --i) {
        double *buf = rb[i & 1];
        if (q == (i & 3)) {
            const double xi = x[i >> 2] * (1.0 / S[i][i]);
            x[i >> 2] = xi;
            buf[c] = xi;
        }
        __syncthreads();
        const double v = buf[c];
#pragma unroll
        for (int k = 0; k < RK; ++k)
            if (4 * k + q < i) x[k] = fma(-S[4 * k + q][i], v, x[k]);
    }
}

// y <- L^{-1} y for the unit lower-triangular L stored strictly below the
// diagonal of S (read only).
__device__ __forceinline__ void solve_unit_lower(Lds *S, double (*rb)[N], double (&y)[RK]) {
    const int c = lane_id(), q = wave_id();
#pragma unroll
    for (int i = 0; i < N; ++i) {
        double *buf = rb[i & 1];
        if (q == (i & 3)) buf[c] = y[i >> 2];
        __syncthreads();
        const double v = buf[c];
#pragma unroll
        for (int k = 0; k < RK; ++k)
            if (4 * k + q > i) y[k] = fma(-S[4 * k + q][i], v, y[k]);
    }
}

// LU without pivoting of M - S with S_j = -sign(current diagonal),
// sign(0) = +: on exit a holds L (unit, strictly below) and U (on and
// above); sg[j] (LDS) the signs.  |U_jj| >= 1 by the sign choice.
__device__ __forceinline__ void lu_sign(double (&a)[RK], double (*rb)[N], double *sg) {
    const int c = lane_id(), q = wave_id();
#pragma unroll
    for (int j = 0; j < N; ++j) {
        double *buf = rb[j & 1];
        if (q == (j & 3)) {
            const double ajj = bcast(a[j >> 2], j);
            const double sj = ajj >= 0.0 ? -1.0 : 1.0;
            if (c == j) a[j >> 2] = ajj - sj;
            buf[c] = a[j >> 2];
            if (c == 0) sg[j] = sj;
        }
        __syncthreads();
        const double u = buf[c];
        const double ip = 1.0 / buf[j];
#pragma unroll
        for (int k = 0; k < RK; ++k)
            if (4 * k + q > j) {
                const double l = bcast(a[k], j) * ip;  // L(r, j)
                a[k] = (c > j) ? fma(-l, u, a[k]) : ((c == j) ? l : a[k]);
            }
    }
}

// Tiles of C = A B (A, B in LDS) on the FP64 MFMA: wave q owns output rows
// 16 q .. 16 q + 15; st(r, c, v) receives every element.  a_upper / b_upper
// skip the k ranges where a triangular operand is zero.
template <typename Store>
__device__ __forceinline__ void matmul(Lds *SA, Lds *SB, bool a_upper, bool b_upper, Store st) {
    const int l = lane_id(), q = wave_id();
    const int r0 = 16 * q;
    v4d acc[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) acc[j] = (v4d){0.0, 0.0, 0.0, 0.0};
    const int kbeg = a_upper ? r0 : 0;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const int kend = b_upper ? 16 * j + 16 : N;
        for (int kk = kbeg; kk < kend; kk += 4) {
            const double av = SA[r0 + (l & 15)][kk + (l >> 4)];
            const double bv = SB[kk + (l >> 4)][16 * j + (l & 15)];
            // swapped operands: lane l, register r holds C(r0 + (l & 15),
            // 16 j + (l >> 4) + 4 r)
            acc[j] = __builtin_amdgcn_mfma_f64_16x16x4f64(bv, av, acc[j], 0, 0, 0);
        }
    }
#pragma unroll
    for (int j = 0; j < 4; ++j)
#pragma unroll
        for (int r = 0; r < 4; ++r) st(r0 + (l & 15), 16 * j + (l >> 4) + 4 * r, acc[j][r]);
}

}  // namespace small
}  // namespace mqr
