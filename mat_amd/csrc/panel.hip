// Column-by-column panel reflectors with their compact-WY T, and the split-K
// partial sums of the block-reflector products.
//
//  k_panel_hh  : QR/__Householder_Blocked.h:62-84 -- b Householder reflectors
//                of an h x b panel, one workgroup, wavefront reductions,
//                followed by their compact-WY T (:86-103) from the Gram
//                matrix V^T V.
#include "mqr_internal.h"
#include "panel_hh.h"
#include "small_dense.h"

namespace mqr {
namespace {

__global__ __launch_bounds__(hh::PNT) void k_panel_hh(int64_t h, int b, double *__restrict__ P,
                                                      int64_t ld, double *__restrict__ tau,
                                                      double *__restrict__ T, int64_t ldt) {
    __shared__ double red[hh::PNT / 64][hh::CH];
    __shared__ double tau_s[small::N];
    __shared__ small::Lds S[small::N];
    __shared__ small::RowBuf rbuf;
    __shared__ double rd[small::N];
    hh::panel_hh_body(h, b, P, ld, tau, T, ldt, hh::PanelLds{red, tau_s, S, &rbuf, rd});
}

}  // namespace

int panel_householder(int64_t h, int b, double *P, int64_t ld, double *tau, double *T,
                      int64_t ldt, hipStream_t s) {
    const double hh = (double)h, bb = b;
    ProfScope ps(K_PANEL_SMALL, 3.0 * hh * bb * bb - 2.0 * bb * bb * bb / 3.0, 24.0 * hh * bb, s);
    hipLaunchKernelGGL(k_panel_hh, dim3(1), dim3(hh::PNT), 0, s, h, b, P, ld, tau, T, ldt);
    MQR_LAUNCH_CHECK();
    return MQR_SUCCESS;
}

// out[0:count] = sum of the splits partials P[s * stride : s * stride + count].
// A workgroup covers 256 / G entries with G threads per entry (thread group
// g sums partials g, g + G, ...); the G sums meet in LDS.  Spreading one
// entry over G threads keeps many loads in flight when there are few entries
// and many partials (a 64 x 64 Gram matrix from 128 row chunks).
__global__ __launch_bounds__(256) void k_reduce_partials(int64_t count, const double *P,
                                                         int splits, int64_t stride, int G,
                                                         double *out, const int *skip) {
    __shared__ double part[256];
    if (skip && *skip) return;  // uniform: the producer found the sum not needed
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
        out[e] = sum;
    }
}

int reduce_partials(int64_t count, double *P, int splits, int64_t stride, hipStream_t s,
                    double *out, const int *skip) {
    if (!out) out = P;
    if (count <= 0 || (splits <= 1 && out == P)) return MQR_SUCCESS;
    ProfScope ps(K_OTHER, (double)count * (splits - 1), 8.0 * count * (splits + 1), s);
    // threads per entry: enough workgroups to spread the loads over the chip
    int G = 1;
    while (G < 16 && G * 2 <= splits && (count * G) / 256 < 512) G *= 2;
    const int E = 256 / G;
    hipLaunchKernelGGL(k_reduce_partials, dim3((unsigned)((count + E - 1) / E)), dim3(256), 0, s,
                       count, (const double *)P, splits, stride, G, out, skip);
    MQR_LAUNCH_CHECK();
    return MQR_SUCCESS;
}

}  // namespace mqr
