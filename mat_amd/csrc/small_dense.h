// Dense kernels for the b x b factors of a panel (b <= 64) in one workgroup
// of 256 threads: Cholesky, LU with sign choice, triangular inverse and
// substitution, and products.
//
// Distributed layout: thread t holds column c = t & 63 of sixteen rows; its
// wave q = t >> 6 owns the row blocks {4 m .. 4 m + 3 : m = 4 k + q}, i.e.
// register j holds row 16 (j >> 2) + 4 q + (j & 3).  Sixteen registers per
// matrix, so several matrices stay resident without spilling.
//
// The eliminations proceed by blocks of four rows.  The wave owning block m
// eliminates inside it with v_readlane broadcasts only (no barrier), then
// publishes the four finished rows through LDS; after one barrier every wave
// applies the rank-4 update to its own later rows.  The LDS row buffer
// alternates between two slots, so one barrier per block suffices: a slot is
// rewritten two blocks later, after every wave has passed the barrier that
// follows its last read.  All barriers are reached by every thread.
//
// Only reciprocals sit on the sequential path: the Cholesky factor is
// computed in LDL^T form (pivots d_i, multipliers a_ir / d_i) and scaled by
// d_i^{-1/2} once at the end, and reciprocals are v_rcp_f64 with a
// second-order correction (recip).  Products of two LDS-resident matrices run
// on the FP64 MFMA.
#pragma once

#include "mqr_internal.h"

namespace mqr {
namespace small {

constexpr int N = 64;    // matrix order (smaller problems are padded with I)
constexpr int NT = 256;  // threads of the workgroup
constexpr int RK = 16;   // rows held per thread
constexpr int BR = 4;    // rows per elimination block
using Lds = double[N + 1];      // row of an LDS matrix S[r][c]
using RowBuf = double[2][BR][N];  // double-buffered block rows

typedef double v4d __attribute__((ext_vector_type(4)));

__device__ __forceinline__ int lane_id() { return threadIdx.x & 63; }
__device__ __forceinline__ int wave_id() { return threadIdx.x >> 6; }
// row held in register j by wave q
__device__ __forceinline__ int drow(int j, int q) { return 16 * (j >> 2) + 4 * q + (j & 3); }

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

// 1 / d: the hardware estimate y0 (relative error below 5e-8 on gfx950)
// corrected to second order, 1 / d = y0 / (1 - e) = y0 (1 + e + e^2 + ...)
// with e = 1 - d y0; the dropped e^3 is below 2^-72, and the result is
// within an ulp.  Three dependent operations after the estimate.
__device__ __forceinline__ double recip(double d) {
    const double y0 = __builtin_amdgcn_rcp(d);
    const double e = fma(-d, y0, 1.0);
    return fma(y0 * e, 1.0 + e, y0);
}

// S = M(0:b, 0:b) (column-major, ld), identity outside; coalesced reads.
// S[r][c] = f(r, c) over the N x N matrix.  Every thread issues its sixteen
// loads before its first LDS store, so the global latency is paid once, not
// once per element.
// A matrix in flight: the sixteen values of the calling thread, element
// e = threadIdx.x + NT it at row e & (N - 1), column e >> 6.  tile_fetch
// issues the loads, tile_stash writes them to LDS; kernels fetch what they
// need early and stash it once the LDS buffer is free.
constexpr int TV = N * N / NT;
template <typename F>
__device__ __forceinline__ void tile_fetch(double (&v)[TV], F f) {
#pragma unroll
    for (int it = 0; it < TV; ++it) {
        const int e = threadIdx.x + it * NT;
        v[it] = f(e & (N - 1), e >> 6);
    }
}
__device__ __forceinline__ void tile_stash(Lds *S, const double (&v)[TV]) {
#pragma unroll
    for (int it = 0; it < TV; ++it) {
        const int e = threadIdx.x + it * NT;
        S[e & (N - 1)][e >> 6] = v[it];
    }
}
template <typename F>
__device__ __forceinline__ void lds_fill(Lds *S, F f) {
    double v[TV];
    tile_fetch(v, f);
    tile_stash(S, v);
}

// Pair layout of a matrix in flight: thread t holds rows 2 (t & 31) + u,
// u = 0, 1, of column (t >> 5) + 8 it as v[2 it + u], so that the two rows
// are one 8- or 16-byte load; eight loads per thread instead of sixteen
// keep a wave's loads of several matrices within its counter of
// outstanding loads.
__device__ __forceinline__ int prow() { return 2 * (int)(threadIdx.x & 31); }
__device__ __forceinline__ int pcol(int it) { return (int)(threadIdx.x >> 5) + 8 * it; }

// v = the calling thread's values of M[r, c] (leading dimension ld) in the
// pair layout; rows r >= rows and columns c >= cols read as zero.  The
// full block with an even leading dimension and an aligned base takes the
// vector loads (the condition is uniform over the workgroup).
template <typename TM>
__device__ __forceinline__ void pair_fetch(double (&v)[TV], const TM *__restrict__ M, int64_t ld,
                                           int64_t rows, int cols) {
    typedef TM v2 __attribute__((ext_vector_type(2)));
    const int r = prow();
    const bool vec = rows >= N && cols >= N && (ld & 1) == 0 &&
                     ((uintptr_t)M & (2 * sizeof(TM) - 1)) == 0;
    if (vec) {
#pragma unroll
        for (int it = 0; it < TV / 2; ++it) {
            const v2 x = *(const v2 *)(M + r + (int64_t)pcol(it) * ld);
            v[2 * it] = (double)x[0];
            v[2 * it + 1] = (double)x[1];
        }
    } else {
#pragma unroll
        for (int it = 0; it < TV / 2; ++it)
#pragma unroll
            for (int u = 0; u < 2; ++u) {
                const int c = pcol(it);
                v[2 * it + u] =
                    (r + u < rows && c < cols) ? (double)M[r + u + (int64_t)c * ld] : 0.0;
            }
    }
}

__device__ __forceinline__ void pair_stash(Lds *S, const double (&v)[TV]) {
#pragma unroll
    for (int it = 0; it < TV / 2; ++it)
#pragma unroll
        for (int u = 0; u < 2; ++u) S[prow() + u][pcol(it)] = v[2 * it + u];
}

// S = M (b x b, ld) padded with the identity to N x N
__device__ __forceinline__ void lds_load(Lds *S, const double *__restrict__ M, int64_t ld, int b) {
    double v[TV];
    pair_fetch(v, M, ld, b, b);
#pragma unroll
    for (int it = 0; it < TV / 2; ++it)
#pragma unroll
        for (int u = 0; u < 2; ++u) {
            const int r = prow() + u, c = pcol(it);
            if (r == c && r >= b) v[2 * it + u] = 1.0;
        }
    pair_stash(S, v);
}

// M(0:rows, 0:cols) = S; coalesced writes, the LDS reads issued first.
__device__ __forceinline__ void lds_store(Lds *S, double *__restrict__ M, int64_t ld, int rows,
                                          int cols) {
    constexpr int IT = N * N / NT;
    double v[IT];
#pragma unroll
    for (int it = 0; it < IT; ++it) {
        const int e = threadIdx.x + it * NT;
        v[it] = S[e & (N - 1)][e >> 6];
    }
#pragma unroll
    for (int it = 0; it < IT; ++it) {
        const int e = threadIdx.x + it * NT, r = e & (N - 1), c = e >> 6;
        if (r < rows && c < cols) M[r + (int64_t)c * ld] = v[it];
    }
}

__device__ __forceinline__ void from_lds(double (&a)[RK], Lds *S) {
    const int c = lane_id(), q = wave_id();
#pragma unroll
    for (int j = 0; j < RK; ++j) a[j] = S[drow(j, q)][c];
}

__device__ __forceinline__ void to_lds(const double (&a)[RK], Lds *S) {
    const int c = lane_id(), q = wave_id();
#pragma unroll
    for (int j = 0; j < RK; ++j) S[drow(j, q)][c] = a[j];
}

__device__ __forceinline__ void set_identity(double (&a)[RK]) {
    const int c = lane_id(), q = wave_id();
#pragma unroll
    for (int j = 0; j < RK; ++j) a[j] = (drow(j, q) == c) ? 1.0 : 0.0;
}

__device__ __forceinline__ void keep_upper(double (&a)[RK]) {
    const int c = lane_id(), q = wave_id();
#pragma unroll
    for (int j = 0; j < RK; ++j) a[j] = (drow(j, q) <= c) ? a[j] : 0.0;
}

// Upper Cholesky factor R (G = R^T R) in place.  Entries below the diagonal
// are updated as well (never read) and zeroed at the end.  *fail (LDS) is set
// when one of the first b pivots is <= its threshold (thr_row[i] when given,
// else thr) or NaN.  dv: 2 N doubles of LDS.
__device__ __forceinline__ void chol(double (&a)[RK], RowBuf &rb, double *dv, int b, double thr,
                                     int *fail, const double *thr_row = nullptr) {
    const int c = lane_id(), q = wave_id();
    double *dinv = dv, *dval = dv + N;
#pragma unroll
    for (int m = 0; m < N / BR; ++m) {
        const int j0 = 4 * (m >> 2), i0 = BR * m;
        double(*raw)[N] = rb[m & 1];
        if (q == (m & 3)) {
#pragma unroll
            for (int s = 0; s < BR; ++s) {
                const int i = i0 + s;
                const double d = bcast(a[j0 + s], i);
                const double di = recip(d);
                if (c == 0) {
                    if (i < b && !(d > (thr_row ? thr_row[i] : thr))) *fail = 1;
                    dinv[i] = di;
                    dval[i] = d;
                }
#pragma unroll
                for (int t = s + 1; t < BR; ++t) {
                    const double l = bcast(a[j0 + s], i0 + t) * di;
                    a[j0 + t] = fma(-l, a[j0 + s], a[j0 + t]);
                }
            }
#pragma unroll
            for (int s = 0; s < BR; ++s) raw[s][c] = a[j0 + s];
        }
        __syncthreads();
        double u[BR], w[BR];
#pragma unroll
        for (int s = 0; s < BR; ++s) {
            u[s] = raw[s][c];
            w[s] = dinv[i0 + s];
        }
#pragma unroll
        for (int j = 0; j < RK; ++j) {
            const int r = drow(j, q);
            if (r >= i0 + BR) {
#pragma unroll
                for (int s = 0; s < BR; ++s) a[j] = fma(-(raw[s][r] * w[s]), u[s], a[j]);
            }
        }
    }
    __syncthreads();
    if (threadIdx.x < N) dinv[threadIdx.x] = 1.0 / sqrt(fmax(dval[threadIdx.x], 0.0));
    __syncthreads();
#pragma unroll
    for (int j = 0; j < RK; ++j) {
        const int r = drow(j, q);
        a[j] = (r <= c) ? a[j] * dinv[r] : 0.0;
    }
}

// x = R^{-1} for the upper-triangular R held in S (read only).  rd: N
// doubles of LDS.
__device__ __forceinline__ void inv_upper(Lds *S, RowBuf &rb, double *rd, double (&x)[RK]) {
    const int c = lane_id(), q = wave_id();
    if (threadIdx.x < N) rd[threadIdx.x] = 1.0 / S[threadIdx.x][threadIdx.x];
    set_identity(x);
    __syncthreads();
#pragma unroll
    for (int m = N / BR - 1; m >= 0; --m) {
        const int j0 = 4 * (m >> 2), i0 = BR * m;
        double(*buf)[N] = rb[m & 1];
        if (q == (m & 3)) {
#pragma unroll
            for (int s = BR - 1; s >= 0; --s) {
                x[j0 + s] *= rd[i0 + s];
#pragma unroll
                for (int t = 0; t < s; ++t) x[j0 + t] = fma(-S[i0 + t][i0 + s], x[j0 + s], x[j0 + t]);
            }
#pragma unroll
            for (int s = 0; s < BR; ++s) buf[s][c] = x[j0 + s];
        }
        __syncthreads();
        double v[BR];
#pragma unroll
        for (int s = 0; s < BR; ++s) v[s] = buf[s][c];
#pragma unroll
        for (int j = 0; j < RK; ++j) {
            const int r = drow(j, q);
            if (r < i0) {
#pragma unroll
                for (int s = 0; s < BR; ++s) x[j] = fma(-S[r][i0 + s], v[s], x[j]);
            }
        }
    }
}

// y <- L^{-1} y for the unit lower-triangular L stored strictly below the
// diagonal of S (read only).
__device__ __forceinline__ void solve_unit_lower(Lds *S, RowBuf &rb, double (&y)[RK]) {
    const int c = lane_id(), q = wave_id();
#pragma unroll
    for (int m = 0; m < N / BR; ++m) {
        const int j0 = 4 * (m >> 2), i0 = BR * m;
        double(*buf)[N] = rb[m & 1];
        if (q == (m & 3)) {
#pragma unroll
            for (int s = 0; s < BR; ++s)
#pragma unroll
                for (int t = s + 1; t < BR; ++t)
                    y[j0 + t] = fma(-S[i0 + t][i0 + s], y[j0 + s], y[j0 + t]);
#pragma unroll
            for (int s = 0; s < BR; ++s) buf[s][c] = y[j0 + s];
        }
        __syncthreads();
        double v[BR];
#pragma unroll
        for (int s = 0; s < BR; ++s) v[s] = buf[s][c];
#pragma unroll
        for (int j = 0; j < RK; ++j) {
            const int r = drow(j, q);
            if (r >= i0 + BR) {
#pragma unroll
                for (int s = 0; s < BR; ++s) y[j] = fma(-S[r][i0 + s], v[s], y[j]);
            }
        }
    }
}

// LU without pivoting of M - S with S_j = -sign(current diagonal),
// sign(0) = +: on exit a holds L (unit, strictly below) and U (on and
// above); sg[j] (LDS) the signs.  |U_jj| >= 1 by the sign choice.  ip: N
// doubles of LDS.
__device__ __forceinline__ void lu_sign(double (&a)[RK], RowBuf &rb, double *ip, double *sg) {
    const int c = lane_id(), q = wave_id();
#pragma unroll
    for (int m = 0; m < N / BR; ++m) {
        const int j0 = 4 * (m >> 2), i0 = BR * m;
        double(*buf)[N] = rb[m & 1];
        if (q == (m & 3)) {
#pragma unroll
            for (int s = 0; s < BR; ++s) {
                const int jp = i0 + s;
                const double ajj = bcast(a[j0 + s], jp);
                const double sj = ajj >= 0.0 ? -1.0 : 1.0;
                const double piv = ajj - sj;
                const double ipv = recip(piv);
                if (c == jp) a[j0 + s] = piv;
                if (c == 0) {
                    sg[jp] = sj;
                    ip[jp] = ipv;
                }
#pragma unroll
                for (int t = s + 1; t < BR; ++t) {
                    const double l = bcast(a[j0 + t], jp) * ipv;
                    a[j0 + t] = (c > jp) ? fma(-l, a[j0 + s], a[j0 + t]) : ((c == jp) ? l : a[j0 + t]);
                }
            }
#pragma unroll
            for (int s = 0; s < BR; ++s) buf[s][c] = a[j0 + s];
        }
        __syncthreads();
        double u[BR], w[BR];
#pragma unroll
        for (int s = 0; s < BR; ++s) {
            u[s] = buf[s][c];
            w[s] = ip[i0 + s];
        }
#pragma unroll
        for (int j = 0; j < RK; ++j) {
            const int r = drow(j, q);
            if (r >= i0 + BR) {
#pragma unroll
                for (int s = 0; s < BR; ++s) {
                    const int jp = i0 + s;
                    const double l = bcast(a[j], jp) * w[s];  // L(r, jp)
                    a[j] = (c > jp) ? fma(-l, u[s], a[j]) : ((c == jp) ? l : a[j]);
                }
            }
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
    for (int kk = kbeg; kk < N; kk += 4) {
        const double av = SA[r0 + (l & 15)][kk + (l >> 4)];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            if (b_upper && kk >= 16 * j + 16) continue;
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
