// LU factorisation with and without partial pivoting, and the direct solve
// built on it (LU(A, L, U), PLU(A, pos, L, U), Solve_direct(A, B):
// LU/__LU.h:18-103, LU/__PLU.h:52-187, LU/__Solve_direct.h:142-220).
//
// Right-looking blocked factorisation in panels of 32 columns on the
// handle's stream:
//   k_getf2      one workgroup factors the h x b panel column by column: the
//                first row of largest magnitude (the reference's and
//                LAPACK's rule) is swapped into place across the panel, the
//                column is scaled by the reciprocal pivot, and the panel's
//                later columns take the rank-1 update;
//   k_laswp      the panel's row swaps applied to the columns left and right
//                of it (one thread per column, swaps in order);
//   k_trsm_lunit U12 = L11^{-1} A12 (one thread per column, L11 in LDS);
//   GEMM         A22 -= L21 U12 on the FP64 MFMA tile.
// The solve permutes B, then runs the unit-lower forward and the upper back
// substitutions blockwise (diagonal blocks by k_trsm_lunit / the QR path's
// upper solve, off-diagonal blocks on the MFMA GEMM).
#include <algorithm>

#include "mqr_internal.h"

namespace mqr {

int trsm_upper_f64(mqr_handle_t hd, bool trans, int64_t n, int64_t nrhs, const double *R,
                   int64_t ldr, double *B, int64_t ldb);

namespace {

constexpr int LNB = 32;   // panel width
constexpr int LNT = 1024; // threads of the panel workgroup
constexpr int LTB = 64;   // diagonal block of the forward substitution

// The h x b panel P (ld) in place; piv[col0 + r] = col0 + the local row
// swapped with row r; *info = the 1-based global column of the first zero
// pivot (atomicMin over the launches; 0 means none was found yet when the
// caller initialised it to INT_MAX).
__global__ __launch_bounds__(LNT) void k_getf2(int64_t h, int b, double *__restrict__ P,
                                               int64_t ld, int pivot, int32_t *__restrict__ piv,
                                               int64_t col0, int32_t *__restrict__ info) {
    __shared__ double red_v[LNT / 64];
    __shared__ int64_t red_i[LNT / 64];
    __shared__ double rowr[LNB];
    __shared__ int64_t pv_row;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    for (int r = 0; r < b; ++r) {
        int64_t p = r;
        if (pivot) {
            // first row of largest magnitude among rows r..h-1
            double bv = -1.0;
            int64_t bi = h;
            for (int64_t i = r + tid; i < h; i += LNT) {
                const double v = fabs(P[i + (int64_t)r * ld]);
                if (v > bv) {
                    bv = v;
                    bi = i;
                }
            }
            for (int o = 32; o > 0; o >>= 1) {
                const double ov = __shfl_xor(bv, o, 64);
                const int64_t oi = __shfl_xor(bi, o, 64);
                if (ov > bv || (ov == bv && oi < bi)) {
                    bv = ov;
                    bi = oi;
                }
            }
            if (lane == 0) {
                red_v[wave] = bv;
                red_i[wave] = bi;
            }
            __syncthreads();
            if (tid == 0) {
                double v = red_v[0];
                int64_t i = red_i[0];
                for (int w = 1; w < LNT / 64; ++w)
                    if (red_v[w] > v || (red_v[w] == v && red_i[w] < i)) {
                        v = red_v[w];
                        i = red_i[w];
                    }
                pv_row = i < h ? i : r;  // every entry NaN: keep the row
            }
            __syncthreads();
            p = pv_row;
            if (tid < b && p != r) {
                const double t = P[r + (int64_t)tid * ld];
                P[r + (int64_t)tid * ld] = P[p + (int64_t)tid * ld];
                P[p + (int64_t)tid * ld] = t;
            }
        }
        if (tid == 0) piv[col0 + r] = (int32_t)(col0 + p);
        __syncthreads();
        if (tid < b) rowr[tid] = P[r + (int64_t)tid * ld];
        __syncthreads();
        const double pvv = rowr[r];
        if (pvv == 0.0) {
            if (tid == 0) atomicMin(info, (int32_t)(col0 + r + 1));
        } else {
            const double inv = 1.0 / pvv;
            for (int64_t i = r + 1 + tid; i < h; i += LNT) {
                double *pi = P + i;
                const double l = pi[(int64_t)r * ld] * inv;
                pi[(int64_t)r * ld] = l;
                for (int c = r + 1; c < b; ++c) pi[(int64_t)c * ld] = fma(-l, rowr[c], pi[(int64_t)c * ld]);
            }
        }
        __syncthreads();
    }
}

// Rows row0 + r and piv[row0 + r] (global row indices) exchanged for
// r = 0..k-1 in order, in each of the ncols columns of A (ld).
__global__ void k_laswp(int64_t ncols, double *__restrict__ A, int64_t lda, int64_t row0, int64_t k,
                        const int32_t *__restrict__ piv) {
    const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= ncols) return;
    double *col = A + c * lda;
    for (int64_t r = 0; r < k; ++r) {
        const int64_t p = piv[row0 + r];
        if (p != row0 + r) {
            const double t = col[row0 + r];
            col[row0 + r] = col[p];
            col[p] = t;
        }
    }
}

// B (b x ncols, ld ldb) = L^{-1} B for the unit lower b x b L (ld ldl),
// b <= 64: one thread per column, L in LDS.
__global__ __launch_bounds__(256) void k_trsm_lunit(int b, const double *__restrict__ L,
                                                    int64_t ldl, double *__restrict__ B,
                                                    int64_t ldb, int64_t ncols) {
    __shared__ double Ls[LTB][LTB + 1];
    for (int e = threadIdx.x; e < b * b; e += 256) {
        const int r = e % b, c = e / b;
        Ls[r][c] = L[r + (int64_t)c * ldl];
    }
    __syncthreads();
    const int64_t col = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (col >= ncols) return;
    double x[LTB];
    double *bc = B + col * ldb;
#pragma unroll
    for (int i = 0; i < LTB; ++i)
        if (i < b) x[i] = bc[i];
#pragma unroll
    for (int i = 1; i < LTB; ++i) {
        if (i < b) {
            double s = x[i];
#pragma unroll
            for (int k = 0; k < LTB; ++k)
                if (k < i) s = fma(-Ls[i][k], x[k], s);
            x[i] = s;
        }
    }
#pragma unroll
    for (int i = 0; i < LTB; ++i)
        if (i < b) bc[i] = x[i];
}

__global__ void k_init_info(int32_t *info) { *info = 0x7fffffff; }
__global__ void k_fix_info(int32_t *info) {
    if (*info == 0x7fffffff) *info = 0;
}

// Forward substitution L X = B for the unit lower n x n L (ld ldl), B (n x
// nrhs, ld ldb) in place: 64-row diagonal blocks, GEMM updates below them.
int trsm_lower_unit(mqr_handle_t hd, int64_t n, int64_t nrhs, const double *L, int64_t ldl,
                    double *B, int64_t ldb) {
    hipStream_t s = hd->stream;
    for (int64_t i0 = 0; i0 < n; i0 += LTB) {
        const int bs = (int)std::min<int64_t>(LTB, n - i0);
        {
            ProfScope ps(K_TRSM, (double)bs * bs * nrhs, 8.0 * (bs * bs + 2.0 * bs * nrhs), s);
            hipLaunchKernelGGL(k_trsm_lunit, dim3((unsigned)((nrhs + 255) / 256)), dim3(256), 0, s,
                               bs, L + i0 + i0 * ldl, ldl, B + i0, ldb, nrhs);
            MQR_LAUNCH_CHECK();
        }
        const int64_t rest = n - i0 - bs;
        if (rest <= 0) continue;
        GemmF64 g;
        g.M = (int)rest; g.N = (int)nrhs; g.K = bs;
        g.A = L + (i0 + bs) + i0 * ldl; g.lda = ldl;
        g.B = B + i0; g.ldb = ldb;
        g.C = B + i0 + bs; g.ldc = ldb;
        g.alpha = -1.0; g.beta = 1.0; g.kind = K_TRSM;
        MQR_CHECK(gemm_f64(g, s));
    }
    return MQR_SUCCESS;
}

}  // namespace

// Blocked LU of A (m x n, lda) in place, with partial pivoting when ipiv
// (min(m, n) entries, device) is given; info (device, may be null): 0 or the
// 1-based column of the first zero pivot.
int getrf_f64(mqr_handle_t hd, int64_t m, int64_t n, double *A, int64_t lda, int32_t *ipiv,
              int32_t *info, int32_t *scratch_piv, int32_t *scratch_info) {
    hipStream_t s = hd->stream;
    const int64_t k = std::min(m, n);
    int32_t *piv = ipiv ? ipiv : scratch_piv;
    int32_t *inf = info ? info : scratch_info;
    hipLaunchKernelGGL(k_init_info, dim3(1), dim3(1), 0, s, inf);
    MQR_LAUNCH_CHECK();
    for (int64_t j = 0; j < k; j += LNB) {
        const int b = (int)std::min<int64_t>(LNB, k - j);
        const int64_t h = m - j;
        {
            ProfScope ps(K_PANEL_SMALL, (double)h * b * b, 8.0 * h * b * b, s);
            hipLaunchKernelGGL(k_getf2, dim3(1), dim3(LNT), 0, s, h, b, A + j + j * lda, lda,
                               ipiv ? 1 : 0, piv, j, inf);
            MQR_LAUNCH_CHECK();
        }
        if (ipiv) {
            // the panel's swaps in the columns left and right of it
            if (j > 0) {
                hipLaunchKernelGGL(k_laswp, dim3((unsigned)((j + 255) / 256)), dim3(256), 0, s, j,
                                   A, lda, j, (int64_t)b, (const int32_t *)piv);
                MQR_LAUNCH_CHECK();
            }
            const int64_t right = n - j - b;
            if (right > 0) {
                hipLaunchKernelGGL(k_laswp, dim3((unsigned)((right + 255) / 256)), dim3(256), 0, s,
                                   right, A + (j + b) * lda, lda, j, (int64_t)b,
                                   (const int32_t *)piv);
                MQR_LAUNCH_CHECK();
            }
        }
        const int64_t right = n - j - b;
        if (right <= 0) continue;
        {
            ProfScope ps(K_TRSM, (double)b * b * right, 8.0 * (b * b + 2.0 * b * right), s);
            hipLaunchKernelGGL(k_trsm_lunit, dim3((unsigned)((right + 255) / 256)), dim3(256), 0, s,
                               b, A + j + j * lda, lda, A + j + (j + b) * lda, lda, right);
            MQR_LAUNCH_CHECK();
        }
        const int64_t below = m - j - b;
        if (below <= 0) continue;
        GemmF64 g;
        g.M = (int)below; g.N = (int)right; g.K = b;
        g.A = A + (j + b) + j * lda; g.lda = lda;
        g.B = A + j + (j + b) * lda; g.ldb = lda;
        g.C = A + (j + b) + (j + b) * lda; g.ldc = lda;
        g.alpha = -1.0; g.beta = 1.0; g.kind = K_UPDATE;
        MQR_CHECK(gemm_f64(g, s));
    }
    hipLaunchKernelGGL(k_fix_info, dim3(1), dim3(1), 0, s, inf);
    MQR_LAUNCH_CHECK();
    return MQR_SUCCESS;
}

// X = A^{-1} B from getrf_f64's factors (LU, ld ldlu; ipiv or none), B (n x
// nrhs, ld ldb) overwritten by X.
int getrs_f64(mqr_handle_t hd, int64_t n, int64_t nrhs, const double *LU, int64_t ldlu,
              const int32_t *ipiv, double *B, int64_t ldb) {
    hipStream_t s = hd->stream;
    if (ipiv) {
        hipLaunchKernelGGL(k_laswp, dim3((unsigned)((nrhs + 255) / 256)), dim3(256), 0, s, nrhs, B,
                           ldb, (int64_t)0, n, ipiv);
        MQR_LAUNCH_CHECK();
    }
    MQR_CHECK(trsm_lower_unit(hd, n, nrhs, LU, ldlu, B, ldb));
    return trsm_upper_f64(hd, false, n, nrhs, LU, ldlu, B, ldb);
}

}  // namespace mqr
