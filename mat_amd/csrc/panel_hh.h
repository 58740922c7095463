// Column-by-column Householder factorisation of an h x b panel (b <= 64)
// with its compact-WY T, by one 256-thread workgroup: the reference's panel
// loop (QR/__Householder_Blocked.h:62-84) with wavefront reductions, and T
// (:86-103) from the Gram matrix V^T V.  The LDS buffers are passed in, so
// that a kernel with its own LDS (the tall-panel reconstruction) can run it
// as its fallback.
#pragma once

#include "mqr_internal.h"
#include "small_dense.h"

namespace mqr {
namespace hh {

constexpr int PNT = 256;  // threads of the panel workgroup
constexpr int CH = 8;     // columns updated per reduction round

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// LDS the factorisation works in
struct PanelLds {
    double (*red)[CH];      // [PNT / 64][CH]
    double *tau_s;          // [small::N]
    small::Lds *S;          // [small::N]
    small::RowBuf *rbuf;
    double *rd;             // [small::N]
};

// P (h x b, ld) in place: R on and above the diagonal, the reflectors below,
// tau[0:b]; T (b x b, ldt) upper triangular.  Every thread of the workgroup
// calls it; it starts and ends at a barrier-consistent point.
__device__ __forceinline__ void panel_hh_body(int64_t h, int b, double *__restrict__ P, int64_t ld,
                                              double *__restrict__ tau, double *__restrict__ T,
                                              int64_t ldt, const PanelLds &L) {
    double (*red)[CH] = L.red;
    double *tau_s = L.tau_s;
    small::Lds *S = L.S;
    small::RowBuf &rbuf = *L.rbuf;
    double *rd = L.rd;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    for (int r = 0; r < b; ++r) {
        double *w = P + (int64_t)r * ld + r;
        const int64_t len = h - r;
        double s = 0.0;
        for (int64_t i = tid; i < len; i += PNT) s = fma(w[i], w[i], s);
        s = wave_sum(s);
        if (lane == 0) red[wave][0] = s;
        __syncthreads();
        double nrm2 = 0.0;
#pragma unroll
        for (int q = 0; q < PNT / 64; ++q) nrm2 += red[q][0];
        double cr = sqrt(nrm2);
        const double alpha = w[0];
        __syncthreads();  // red[] and w[0] read by everyone
        if (cr == 0.0) {
            if (tid == 0) tau[r] = tau_s[r] = 0.0;
            continue;
        }
        if (alpha < 0.0) cr = -cr;
        const double w0 = alpha + cr;
        const double hr = 1.0 / w0;
        const double t = w0 / cr;
        for (int64_t i = 1 + tid; i < len; i += PNT) w[i] = w[i] * hr;
        if (tid == 0) {
            w[0] = -cr;
            tau[r] = tau_s[r] = t;
        }
        __syncthreads();
        for (int c0 = r + 1; c0 < b; c0 += CH) {
            const int nc = min(CH, b - c0);
            double d[CH], top[CH];
#pragma unroll
            for (int q = 0; q < CH; ++q) {
                d[q] = 0.0;
                top[q] = q < nc ? P[(int64_t)(c0 + q) * ld + r] : 0.0;
            }
            for (int64_t i = 1 + tid; i < len; i += PNT) {
                const double vi = w[i];
#pragma unroll
                for (int q = 0; q < CH; ++q)
                    if (q < nc) d[q] = fma(vi, P[(int64_t)(c0 + q) * ld + r + i], d[q]);
            }
#pragma unroll
            for (int q = 0; q < CH; ++q) {
                const double sq = wave_sum(d[q]);
                if (lane == 0) red[wave][q] = sq;
            }
            __syncthreads();
            double coef[CH];
#pragma unroll
            for (int q = 0; q < CH; ++q) {
                double tot = top[q];
#pragma unroll
                for (int u = 0; u < PNT / 64; ++u) tot += red[u][q];
                coef[q] = -tot * t;
            }
            __syncthreads();
            for (int64_t i = tid; i < len; i += PNT) {
                const double vi = i == 0 ? 1.0 : w[i];
#pragma unroll
                for (int q = 0; q < CH; ++q)
                    if (q < nc) {
                        double *p = P + (int64_t)(c0 + q) * ld + r + i;
                        *p = fma(vi, coef[q], *p);
                    }
            }
        }
        __syncthreads();
    }
    // T: the forward columnwise recurrence T(0:i, i) = -tau_i T(0:i, 0:i)
    // V(:, 0:i)^T v_i, T(i, i) = tau_i (QR/__Householder_Blocked.h:86-103) has
    // the closed form T^{-1} = strict_upper(V^T V) + diag(1 / tau).  G = V^T V
    // is accumulated over 64-row slices of V staged in LDS; a reflector with
    // tau = 0 is the identity, its row and column are removed from the
    // inverse and left zero in T, which is what the recurrence produces.
    const int c = lane, q = wave;
    double g[small::RK];
#pragma unroll
    for (int k = 0; k < small::RK; ++k) g[k] = 0.0;
    for (int64_t r0 = 0; r0 < h; r0 += small::N) {
        __syncthreads();
        for (int e = tid; e < small::N * small::N; e += PNT) {
            const int rr = e & (small::N - 1), cc = e >> 6;
            const int64_t r = r0 + rr;
            double v = 0.0;
            if (cc < b && r < h) v = r > cc ? P[r + (int64_t)cc * ld] : (r == cc ? 1.0 : 0.0);
            S[rr][cc] = v;
        }
        __syncthreads();
        for (int rr = 0; rr < small::N; ++rr) {
            const double vc = S[rr][c];
#pragma unroll
            for (int j = 0; j < small::RK; ++j) g[j] = fma(S[rr][small::drow(j, q)], vc, g[j]);
        }
    }
    __syncthreads();
    const double tc = c < b ? tau_s[c] : 1.0;
#pragma unroll
    for (int j = 0; j < small::RK; ++j) {
        const int r = small::drow(j, q);
        const double tr = r < b ? tau_s[r] : 1.0;
        double v = 0.0;
        if (r < c && c < b) v = (tr == 0.0 || tc == 0.0) ? 0.0 : g[j];
        else if (r == c) v = tc != 0.0 ? 1.0 / tc : 1.0;
        g[j] = v;
    }
    small::to_lds(g, S);
    __syncthreads();
    small::inv_upper(S, rbuf, rd, g);
#pragma unroll
    for (int j = 0; j < small::RK; ++j) {
        const int r = small::drow(j, q);
        if (r < b && c < b) {
            double v = r <= c ? g[j] : 0.0;
            if (r == c && tc == 0.0) v = 0.0;
            T[r + (int64_t)c * ldt] = v;
        }
    }
}

}  // namespace hh
}  // namespace mqr
