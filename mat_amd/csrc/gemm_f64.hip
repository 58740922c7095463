// FP64 MFMA GEMM tile for gfx950: the device counterpart of the reference's
// packed MR x NR micro-kernel and its macro-kernel drivers
// (Mul/__micro_kernel_packed_double.h:301-397,763-817; Mul/Mul.h:239-322).
//
// Instead of packing into MR/NR panels in a separate pass
// (Mul/__pack_double.h), each workgroup stages a 64 x 16 slice of op(A) and a
// 16 x 64 slice of op(B) in LDS, applying the unit-lower-trapezoid mask of a
// reflector block V on the fly (the reference's __pack_L_block /
// __pack_LT_block, Mul/__pack.h:108-189).  Four waves each own a 32 x 32
// quarter of the 64 x 64 output tile as 2 x 2 v_mfma_f64_16x16x4_f64 blocks.
#include <climits>

#include "mqr_internal.h"

namespace mqr {
namespace {

typedef double v4d __attribute__((ext_vector_type(4)));

constexpr int TBM = 64, TBN = 64, TBK = 16, NT = 256;

template <int MASK>
__device__ __forceinline__ double ld_masked(const double *A, int64_t ld, int r, int c,
                                            int mask_cols) {
    if (MASK == MASK_VLOWER && c < mask_cols) {
        if (r < c) return 0.0;
        if (r == c) return 1.0;
    }
    return A[(int64_t)r + (int64_t)c * ld];
}

template <bool TA, bool TB, int AM, int BMK>
__global__ __launch_bounds__(NT) void k_gemm_f64(GemmF64 g, int kchunk) {
    __shared__ double As[TBK][TBM + 1];
    __shared__ double Bs[TBK][TBN + 1];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = wave & 1, wn = wave >> 1;
    const int m0 = blockIdx.x * TBM, n0 = blockIdx.y * TBN;
    const int kb = blockIdx.z * kchunk;
    const int ke = min(g.K, kb + kchunk);

    v4d acc[2][2];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) acc[i][j] = (v4d){0.0, 0.0, 0.0, 0.0};

    for (int k0 = kb; k0 < ke; k0 += TBK) {
#pragma unroll
        for (int t = 0; t < (TBM * TBK) / NT; ++t) {
            const int e = tid + t * NT;
            int mm, kk;
            if (TA) { kk = e % TBK; mm = e / TBK; }
            else { mm = e % TBM; kk = e / TBM; }
            const int gm = m0 + mm, gk = k0 + kk;
            double v = 0.0;
            if (gm < g.M && gk < ke)
                v = TA ? ld_masked<AM>(g.A, g.lda, gk, gm, INT_MAX)
                       : ld_masked<AM>(g.A, g.lda, gm, gk, INT_MAX);
            As[kk][mm] = v;
        }
#pragma unroll
        for (int t = 0; t < (TBN * TBK) / NT; ++t) {
            const int e = tid + t * NT;
            int nn, kk;
            if (!TB) { kk = e % TBK; nn = e / TBK; }
            else { nn = e % TBN; kk = e / TBN; }
            const int gn = n0 + nn, gk = k0 + kk;
            double v = 0.0;
            if (gn < g.N && gk < ke)
                v = !TB ? ld_masked<BMK>(g.B, g.ldb, gk, gn, g.bmask_cols)
                        : ld_masked<BMK>(g.B, g.ldb, gn, gk, g.bmask_cols);
            Bs[kk][nn] = v;
        }
        __syncthreads();
#pragma unroll
        for (int kk = 0; kk < TBK; kk += 4) {
            double a[2], b[2];
#pragma unroll
            for (int i = 0; i < 2; ++i) a[i] = As[kk + (lane >> 4)][wm * 32 + i * 16 + (lane & 15)];
#pragma unroll
            for (int j = 0; j < 2; ++j) b[j] = Bs[kk + (lane >> 4)][wn * 32 + j * 16 + (lane & 15)];
            // operands swapped so that the 16 lanes of a row group hold 16
            // consecutive rows of one output column (coalesced column-major stores)
#pragma unroll
            for (int i = 0; i < 2; ++i)
#pragma unroll
                for (int j = 0; j < 2; ++j)
                    acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(b[j], a[i], acc[i][j], 0, 0, 0);
        }
        __syncthreads();
    }

    double *Cout = g.C + (g.splits > 1 ? (int64_t)blockIdx.z * g.split_stride : 0);
    const bool direct = g.splits > 1 || g.beta == 0.0;
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                // v_mfma_f64_16x16x4_f64: lane l, accumulator r holds output
                // row (l >> 4) + 4 r, column l & 15 (of the swapped product)
                const int gm = m0 + wm * 32 + i * 16 + (lane & 15);
                const int gn = n0 + wn * 32 + j * 16 + (lane >> 4) + 4 * r;
                if (gm < g.M && gn < g.N) {
                    const int64_t idx = (int64_t)gm + (int64_t)gn * g.ldc;
                    const double v = g.alpha * acc[i][j][r];
                    Cout[idx] = direct ? v : g.beta * Cout[idx] + v;
                }
            }
}

template <bool TA, bool TB, int AM, int BMK>
void launch(const GemmF64 &g, hipStream_t s) {
    int kchunk = (g.K + g.splits - 1) / g.splits;
    kchunk = ((kchunk + TBK - 1) / TBK) * TBK;
    dim3 grid((g.M + TBM - 1) / TBM, (g.N + TBN - 1) / TBN, g.splits);
    hipLaunchKernelGGL((k_gemm_f64<TA, TB, AM, BMK>), grid, dim3(NT), 0, s, g, kchunk);
}

__global__ void k_scale(int64_t m, int64_t n, double beta, double *C, int64_t ldc) {
    const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= m * n) return;
    double *p = C + idx % m + (idx / m) * ldc;
    *p = beta == 0.0 ? 0.0 : beta * *p;
}

}  // namespace

int scale_f64(int64_t m, int64_t n, double beta, double *C, int64_t ldc, hipStream_t s) {
    if (m <= 0 || n <= 0) return MQR_SUCCESS;
    hipLaunchKernelGGL(k_scale, dim3((unsigned)((m * n + 255) / 256)), dim3(256), 0, s, m, n, beta,
                       C, ldc);
    MQR_LAUNCH_CHECK();
    return MQR_SUCCESS;
}

int gemm_pick_splits(int M, int N, int K, int target_blocks) {
    const int tiles = ((M + TBM - 1) / TBM) * ((N + TBN - 1) / TBN);
    if (tiles >= target_blocks || K <= 512) return 1;
    int s = (target_blocks + tiles - 1) / tiles;
    const int maxs = (K + 255) / 256;
    if (s > maxs) s = maxs;
    if (s > 64) s = 64;
    return s < 1 ? 1 : s;
}

int gemm_f64(const GemmF64 &g, hipStream_t s) {
    if (g.M <= 0 || g.N <= 0) return MQR_SUCCESS;
    if (g.K <= 0) return MQR_ERR_INTERNAL;  // callers never ask for an empty sum
    const int key = (g.ta ? 8 : 0) | (g.tb ? 4 : 0) | (g.amask ? 2 : 0) | (g.bmask ? 1 : 0);
    const double M = g.M, N = g.N, K = g.K;
    const double out_rw = (g.splits > 1 || g.beta == 0.0) ? 1.0 : 2.0;
    ProfScope ps(g.kind, 2.0 * M * N * K, 8.0 * (M * K + K * N + out_rw * M * N * g.splits), s);
    switch (key) {
        case 0: launch<false, false, MASK_NONE, MASK_NONE>(g, s); break;
        case 2: launch<false, false, MASK_VLOWER, MASK_NONE>(g, s); break;
        case 8: launch<true, false, MASK_NONE, MASK_NONE>(g, s); break;
        case 10: launch<true, false, MASK_VLOWER, MASK_NONE>(g, s); break;
        case 11: launch<true, false, MASK_VLOWER, MASK_VLOWER>(g, s); break;
        case 4: launch<false, true, MASK_NONE, MASK_NONE>(g, s); break;
        case 12: launch<true, true, MASK_NONE, MASK_NONE>(g, s); break;
        default: return MQR_ERR_UNSUPPORTED;
    }
    MQR_LAUNCH_CHECK();
    return MQR_SUCCESS;
}

}  // namespace mqr
