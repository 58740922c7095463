// Panel reflectors, compact-WY T and the T-product of the block reflector.
//
//  k_panel_hh  : QR/__Householder_Blocked.h:62-84 -- b Householder reflectors
//                of an h x b panel, one workgroup, wavefront reductions.
//  k_larft     : QR/__Householder_Blocked.h:86-103 -- the forward columnwise
//                T recurrence T(0:i,i) = -tau_i T(0:i,0:i) s, with
//                s_l = V(:,l)^T V(:,i) taken from the Gram matrix V^T V that
//                the MFMA GEMM produced.
//  k_block_w2  : the middle product of the trailing update,
//                W2 = T^T W (QR/__Householder_Blocked.h:180-190) or T W
//                (Q formation, :299-308), with W summed from split-K partials.
#include "mqr_internal.h"

namespace mqr {
namespace {

constexpr int PNT = 256;  // threads of the panel workgroup
constexpr int CH = 8;     // columns updated per reduction round

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

__global__ __launch_bounds__(PNT) void k_panel_hh(int64_t h, int b, double *__restrict__ P,
                                                  int64_t ld, double *__restrict__ tau) {
    __shared__ double red[PNT / 64][CH];
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
            if (tid == 0) tau[r] = 0.0;
            continue;
        }
        if (alpha < 0.0) cr = -cr;
        const double w0 = alpha + cr;
        const double hr = 1.0 / w0;
        const double t = w0 / cr;
        for (int64_t i = 1 + tid; i < len; i += PNT) w[i] = w[i] * hr;
        if (tid == 0) {
            w[0] = -cr;
            tau[r] = t;
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
}

// T from the split-K partial Gram matrices of V (slot s at Gp + s * pstride,
// ld = b).  Dynamic LDS: 2 * b * b doubles (b <= 64).
__global__ __launch_bounds__(256) void k_larft(int b, const double *__restrict__ Gp, int splits,
                                               int64_t pstride, const double *__restrict__ tau,
                                               double *__restrict__ T, int64_t ldt) {
    extern __shared__ double sm[];
    double *G = sm;
    double *Tl = sm + b * b;
    const int tid = threadIdx.x;
    for (int e = tid; e < b * b; e += blockDim.x) {
        double acc = 0.0;
        for (int s = 0; s < splits; ++s) acc += Gp[(int64_t)s * pstride + e];
        G[e] = acc;
        Tl[e] = 0.0;
    }
    __syncthreads();
    if (tid == 0) Tl[0] = tau[0];
    __syncthreads();
    for (int i = 1; i < b; ++i) {
        const double ti = tau[i];
        if (tid < i) {
            double acc = 0.0;
            for (int c = tid; c < i; ++c) acc = fma(Tl[tid + c * b], G[c + i * b], acc);
            Tl[tid + i * b] = -ti * acc;
        }
        if (tid == 0) Tl[i + i * b] = ti;
        __syncthreads();
    }
    for (int e = tid; e < b * b; e += blockDim.x) {
        const int r = e % b, c = e / b;
        T[r + (int64_t)c * ldt] = r <= c ? Tl[e] : 0.0;
    }
}

constexpr int W2COLS = 32;

// W2(:, n) = op(T) * sum_s P_s(:, n), for 32 columns per workgroup.
__global__ __launch_bounds__(256) void k_block_w2(int b, int64_t nc, const double *__restrict__ T,
                                                  int64_t ldt, int trans,
                                                  const double *__restrict__ P, int splits,
                                                  int64_t pstride, double *__restrict__ W2) {
    extern __shared__ double sm[];
    double *Ts = sm;          // b * b
    double *Ws = sm + b * b;  // b * W2COLS
    const int tid = threadIdx.x;
    const int64_t n0 = (int64_t)blockIdx.x * W2COLS;
    for (int e = tid; e < b * b; e += blockDim.x) {
        const int r = e % b, c = e / b;
        Ts[e] = r <= c ? T[r + (int64_t)c * ldt] : 0.0;
    }
    for (int e = tid; e < b * W2COLS; e += blockDim.x) {
        const int r = e % b;
        const int64_t n = n0 + e / b;
        double acc = 0.0;
        if (n < nc)
            for (int s = 0; s < splits; ++s) acc += P[(int64_t)s * pstride + r + n * b];
        Ws[e] = acc;
    }
    __syncthreads();
    for (int e = tid; e < b * W2COLS; e += blockDim.x) {
        const int r = e % b, col = e / b;
        const int64_t n = n0 + col;
        if (n >= nc) continue;
        double acc = 0.0;
        if (trans) {
            for (int l = 0; l <= r; ++l) acc = fma(Ts[l + r * b], Ws[l + col * b], acc);
        } else {
            for (int l = r; l < b; ++l) acc = fma(Ts[r + l * b], Ws[l + col * b], acc);
        }
        W2[r + n * b] = acc;
    }
}

bool g_lds_set = false;

int ensure_lds_attrs() {
    if (g_lds_set) return MQR_SUCCESS;
    MQR_HIP(hipFuncSetAttribute((const void *)k_larft, hipFuncAttributeMaxDynamicSharedMemorySize,
                                2 * kMaxNb * kMaxNb * 8));
    MQR_HIP(hipFuncSetAttribute((const void *)k_block_w2,
                                hipFuncAttributeMaxDynamicSharedMemorySize,
                                (kMaxNb * kMaxNb + kMaxNb * W2COLS) * 8));
    g_lds_set = true;
    return MQR_SUCCESS;
}

}  // namespace

int panel_householder(int64_t h, int b, double *P, int64_t ld, double *tau, hipStream_t s) {
    const double hh = (double)h, bb = b;
    ProfScope ps(K_PANEL, 2.0 * hh * bb * bb - 2.0 * bb * bb * bb / 3.0, 16.0 * hh * bb, s);
    hipLaunchKernelGGL(k_panel_hh, dim3(1), dim3(PNT), 0, s, h, b, P, ld, tau);
    MQR_LAUNCH_CHECK();
    return MQR_SUCCESS;
}

int larft_from_gram(int b, const double *Gpart, int splits, int64_t pstride, const double *tau,
                    double *T, int64_t ldt, hipStream_t s) {
    MQR_CHECK(ensure_lds_attrs());
    const size_t lds = (size_t)2 * b * b * sizeof(double);
    ProfScope ps(K_LARFT, (double)b * b * b / 3.0, 8.0 * b * b * (splits + 1), s);
    hipLaunchKernelGGL(k_larft, dim3(1), dim3(256), lds, s, b, Gpart, splits, pstride, tau, T,
                       ldt);
    MQR_LAUNCH_CHECK();
    return MQR_SUCCESS;
}

int block_w2(int b, int64_t nc, const double *T, int64_t ldt, bool trans, const double *P,
             int splits, int64_t pstride, double *W2, hipStream_t s) {
    if (nc <= 0) return MQR_SUCCESS;
    MQR_CHECK(ensure_lds_attrs());
    const size_t lds = (size_t)(b * b + b * W2COLS) * sizeof(double);
    const unsigned blocks = (unsigned)((nc + W2COLS - 1) / W2COLS);
    ProfScope ps(K_W2, (double)b * b * nc, 8.0 * b * nc * (splits + 1), s);
    hipLaunchKernelGGL(k_block_w2, dim3(blocks), dim3(256), lds, s, b, nc, T, ldt, trans ? 1 : 0,
                       P, splits, pstride, W2);
    MQR_LAUNCH_CHECK();
    return MQR_SUCCESS;
}

}  // namespace mqr
