// Column-by-column panel reflectors with their compact-WY T, and the split-K
// partial sums of the block-reflector products.
//
//  k_panel_hh  : QR/__Householder_Blocked.h:62-84 -- b Householder reflectors
//                of an h x b panel, one workgroup, wavefront reductions,
//                followed by their compact-WY T (:86-103) from the Gram
//                matrix V^T V.
#include "mqr_internal.h"
#include "small_dense.h"

namespace mqr {
namespace {

constexpr int PNT = 256;  // threads of the panel workgroup
constexpr int CH = 8;     // columns updated per reduction round

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// gate: when non-null, the kernel only runs if *gate is nonzero (the fast
// panel raised its fallback flag).
__global__ __launch_bounds__(PNT) void k_panel_hh(int64_t h, int b, double *__restrict__ P,
                                                  int64_t ld, double *__restrict__ tau,
                                                  double *__restrict__ T, int64_t ldt,
                                                  const int *__restrict__ gate) {
    __shared__ double red[PNT / 64][CH];
    __shared__ double tau_s[small::N];
    __shared__ small::Lds S[small::N];
    __shared__ small::RowBuf rbuf;
    __shared__ double rd[small::N];
    if (gate && *gate == 0) return;
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

}  // namespace

int panel_householder(int64_t h, int b, double *P, int64_t ld, double *tau, double *T,
                      int64_t ldt, hipStream_t s) {
    const double hh = (double)h, bb = b;
    ProfScope ps(K_PANEL_SMALL, 3.0 * hh * bb * bb - 2.0 * bb * bb * bb / 3.0, 24.0 * hh * bb, s);
    hipLaunchKernelGGL(k_panel_hh, dim3(1), dim3(PNT), 0, s, h, b, P, ld, tau, T, ldt,
                       (const int *)nullptr);
    MQR_LAUNCH_CHECK();
    return MQR_SUCCESS;
}

int panel_householder_gated(int64_t h, int b, double *P, int64_t ld, double *tau, double *T,
                            int64_t ldt, const int *gate, hipStream_t s) {
    hipLaunchKernelGGL(k_panel_hh, dim3(1), dim3(PNT), 0, s, h, b, P, ld, tau, T, ldt, gate);
    MQR_LAUNCH_CHECK();
    return MQR_SUCCESS;
}

// P[0:count] = sum over s of P[s * stride + 0:count]
// P[0:count] = sum of the splits partials P[s * stride : s * stride + count].
// A workgroup covers 256 / G entries with G threads per entry (thread group
// g sums partials g, g + G, ...); the G sums meet in LDS.  Spreading one
// entry over G threads keeps many loads in flight when there are few entries
// and many partials (a 64 x 64 Gram matrix from 128 row chunks).
__global__ __launch_bounds__(256) void k_reduce_partials(int64_t count, double *__restrict__ P,
                                                         int splits, int64_t stride, int G) {
    __shared__ double part[256];
    const int E = 256 / G;
    const int t = threadIdx.x, g = t / E;
    const int64_t e = (int64_t)blockIdx.x * E + (t % E);
    double a0 = 0.0, a1 = 0.0;
    if (e < count) {
        int s = g;
        for (; s + G < splits; s += 2 * G) {
            a0 += P[(int64_t)s * stride + e];
            a1 += P[(int64_t)(s + G) * stride + e];
        }
        if (s < splits) a0 += P[(int64_t)s * stride + e];
    }
    part[t] = a0 + a1;
    __syncthreads();
    if (g == 0 && e < count) {
        double sum = part[t];
        for (int k = 1; k < G; ++k) sum += part[t + k * E];
        P[e] = sum;
    }
}

int reduce_partials(int64_t count, double *P, int splits, int64_t stride, hipStream_t s) {
    if (splits <= 1 || count <= 0) return MQR_SUCCESS;
    ProfScope ps(K_OTHER, (double)count * (splits - 1), 8.0 * count * (splits + 1), s);
    // threads per entry: enough workgroups to spread the loads over the chip
    int G = 1;
    while (G < 16 && G * 2 <= splits && (count * G) / 256 < 512) G *= 2;
    const int E = 256 / G;
    hipLaunchKernelGGL(k_reduce_partials, dim3((unsigned)((count + E - 1) / E)), dim3(256), 0, s,
                       count, P, splits, stride, G);
    MQR_LAUNCH_CHECK();
    return MQR_SUCCESS;
}

}  // namespace mqr
