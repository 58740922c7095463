// QR adjoint (reverse-mode derivative of the thin QR A = Q R, m >= n), for a
// strided batch of equally sized matrices in FP32 or FP64.  No reference
// counterpart; built from the same MFMA GEMM tiles as the factorisation.
//
//   M  = R dR^T - dQ^T Q                       (n x n)
//   dA = [dQ + Q copyltu(M)] R^{-T},  copyltu(M) = tril(M) + tril(M, -1)^T
//
// R^{-T} is applied as a blocked right triangular solve X R^T = B over
// 64-column blocks from the last: X_J = B_J R_JJ^{-T} (the 64 x 64 diagonal
// inverses are formed in FP64 in one workgroup each), then
// B_<J -= X_J R_<J,J^T; every product is a batched MFMA GEMM.  B is formed
// in a workspace and X_J is written straight into dA, so no product reads
// and writes the same block.
#include <algorithm>

#include "mqr_internal.h"
#include "small_dense.h"

namespace mqr {
namespace {

constexpr int AB = small::N;  // block of the triangular solve

// One batched GEMM on the FP32 or FP64 tile: C = alpha op(A) op(B) + beta Cin,
// with Cin = C unless given.
int bgemm(bool f64, bool ta, bool tb, int M, int N, int K, double alpha, const void *A,
          int64_t lda, int64_t sA, const void *B, int64_t ldb, int64_t sB, double beta, void *C,
          int64_t ldc, int64_t sC, int batch, int kind, hipStream_t s,
          const void *Cin = nullptr, int64_t ldcin = 0, int64_t sCin = 0) {
    if (M <= 0 || N <= 0 || K <= 0) return MQR_SUCCESS;
    if (f64) {
        GemmF64 g;
        g.M = M; g.N = N; g.K = K; g.ta = ta; g.tb = tb;
        g.A = (const double *)A; g.lda = lda; g.B = (const double *)B; g.ldb = ldb;
        g.C = (double *)C; g.ldc = ldc; g.alpha = alpha; g.beta = beta;
        g.Cin = (const double *)Cin; g.ldcin = ldcin; g.strideCin = sCin;
        g.batch = batch; g.strideA = sA; g.strideB = sB; g.strideC = sC; g.kind = kind;
        return gemm_f64(g, s);
    }
    GemmF32 g;
    g.M = M; g.N = N; g.K = K; g.ta = ta; g.tb = tb;
    g.A = (const float *)A; g.lda = lda; g.B = (const float *)B; g.ldb = ldb;
    g.C = (float *)C; g.ldc = ldc; g.alpha = (float)alpha; g.beta = (float)beta;
    g.Cin = (const float *)Cin; g.ldcin = ldcin; g.strideCin = sCin;
    g.batch = batch; g.strideA = sA; g.strideB = sB; g.strideC = sC; g.kind = kind;
    return gemm_f32(g, s);
}

// Y_b (rows x cols, ldy) = X_b (ldx), or zero when X is null; grid.y = batch
template <typename T>
__global__ void k_copy2d(int64_t rows, int64_t cols, const T *__restrict__ X, int64_t ldx,
                         int64_t sx, T *__restrict__ Y, int64_t ldy, int64_t sy) {
    const int64_t b = blockIdx.y;
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < rows * cols;
         e += (int64_t)gridDim.x * blockDim.x) {
        const int64_t r = e % rows, c = e / rows;
        Y[b * sy + r + c * ldy] = X ? X[b * sx + r + c * ldx] : T(0);
    }
}

// Y_b (n x n, ld n) = upper triangle of X_b (ldx), zero below the diagonal
template <typename T>
__global__ void k_triu(int64_t n, const T *__restrict__ X, int64_t ldx, int64_t sx,
                       T *__restrict__ Y, int64_t sy) {
    const int64_t b = blockIdx.y;
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n * n;
         e += (int64_t)gridDim.x * blockDim.x) {
        const int64_t r = e % n, c = e / n;
        Y[b * sy + r + c * n] = r <= c ? X[b * sx + r + c * ldx] : T(0);
    }
}

// M_b <- copyltu(M_b) = tril(M_b) + tril(M_b, -1)^T (n x n, ld n), in place,
// by 32 x 32 tiles through LDS: workgroup t of grid.x takes the tile at
// block row I >= block column J (t = I (I + 1) / 2 + J), reads it with
// column-contiguous loads and writes its transpose into tile (J, I) above
// the diagonal, so both sides move whole column pieces.  Only tiles on or
// below the diagonal are read and only entries above it written, so the
// workgroups never touch each other's data.
constexpr int CL_T = 32;
template <typename T>
__global__ __launch_bounds__(256) void k_copyltu(int64_t n, T *__restrict__ M, int64_t sm) {
    __shared__ T tile[CL_T][CL_T + 1];  // tile[local column][local row]
    const int64_t t = blockIdx.x;
    int64_t I = (int64_t)((sqrt(8.0 * (double)t + 1.0) - 1.0) * 0.5);
    while (I * (I + 1) / 2 > t) --I;
    while ((I + 1) * (I + 2) / 2 <= t) ++I;
    const int64_t J = t - I * (I + 1) / 2;
    T *Mb = M + (int64_t)blockIdx.y * sm;
    const int tx = threadIdx.x & (CL_T - 1), ty = threadIdx.x / CL_T;  // 32 x 8
    for (int k = 0; k < CL_T; k += 256 / CL_T) {
        const int64_t r = I * CL_T + tx, c = J * CL_T + ty + k;
        if (r < n && c < n) tile[ty + k][tx] = Mb[r + c * n];
    }
    __syncthreads();
    // (r2, c2) above the diagonal takes M(c2, r2) = tile[r2 - J 32][c2 - I 32]
    for (int k = 0; k < CL_T; k += 256 / CL_T) {
        const int64_t r2 = J * CL_T + tx, c2 = I * CL_T + ty + k;
        if (r2 < n && c2 < n && r2 < c2) Mb[r2 + c2 * n] = tile[tx][ty + k];
    }
}

// Rinv[(b * nblk + J) * AB * AB] (AB x AB, ld AB) = inverse of the J-th
// diagonal block of the upper R_b, computed in FP64; grid = (nblk, batch)
template <typename T>
__global__ __launch_bounds__(small::NT) void k_inv_diag_batched(int64_t n, const T *__restrict__ R,
                                                                int64_t ldr, int64_t sr,
                                                                T *__restrict__ Rinv) {
    __shared__ small::Lds S[small::N];
    __shared__ small::RowBuf rb;
    __shared__ double rd[small::N];
    const int J = blockIdx.x, nblk = gridDim.x;
    const int64_t b = blockIdx.y;
    const int64_t i0 = (int64_t)J * AB;
    const int bj = (int)std::min<int64_t>(AB, n - i0);
    const T *Rb = R + b * sr;
    for (int e = threadIdx.x; e < AB * AB; e += small::NT) {
        const int r = e & (AB - 1), c = e >> 6;
        S[r][c] = (r < bj && c < bj) ? (r <= c ? (double)Rb[i0 + r + (i0 + c) * ldr] : 0.0)
                                     : (r == c ? 1.0 : 0.0);
    }
    __syncthreads();
    double x[small::RK];
    small::inv_upper(S, rb, rd, x);
    __syncthreads();
    small::to_lds(x, S);
    __syncthreads();
    T *out = Rinv + (b * nblk + J) * (int64_t)AB * AB;
    for (int e = threadIdx.x; e < AB * AB; e += small::NT) {
        const int r = e & (AB - 1), c = e >> 6;
        out[r + c * AB] = (T)S[r][c];
    }
}

template <typename T>
int qr_backward_t(mqr_handle_t hd, int64_t batch, int64_t m, int64_t n, const T *Q, int64_t ldq,
                  int64_t sq, const T *R, int64_t ldr, int64_t sr, const T *dQ, int64_t lddq,
                  int64_t sdq, const T *dR, int64_t lddr, int64_t sdr, T *dA, int64_t ldda,
                  int64_t sda) {
    const bool f64 = sizeof(T) == 8;
    hipStream_t s = hd->stream;
    const int nblk = (int)((n + AB - 1) / AB);
    const int B = (int)batch;
    // workspace: M (n x n), triu(dR) (n x n), the diagonal inverses,
    // B (m x n), per matrix
    const int64_t sM = n * n, sI = (int64_t)nblk * AB * AB, sW = m * n;
    MQR_CHECK(ws_reserve(hd, hd->ws, sizeof(T) * (size_t)batch * (2 * sM + sI + sW) + 4096));
    T *Mw = (T *)hd->ws.ptr;
    T *Uw = Mw + batch * sM;
    T *Iw = Uw + batch * sM;
    T *Ww = Iw + batch * sI;
    const unsigned gx = (unsigned)std::min<int64_t>(1024, (n * n + 255) / 256);
    // M = R triu(dR)^T - dQ^T Q: R's strictly lower part is identically zero,
    // so only dR's upper triangle carries a derivative
    if (dR) {
        {
            ProfScope ps(K_OTHER, 0.0, 16.0 * batch * n * n, s);
            hipLaunchKernelGGL(k_triu<T>, dim3(gx, B), dim3(256), 0, s, n, dR, lddr, sdr, Uw, sM);
            MQR_LAUNCH_CHECK();
        }
        MQR_CHECK(bgemm(f64, false, true, (int)n, (int)n, (int)n, 1.0, R, ldr, sr, Uw, n, sM,
                        0.0, Mw, n, sM, B, K_GEMM, s));
    } else {
        ProfScope ps(K_OTHER, 0.0, 8.0 * batch * n * n, s);
        hipLaunchKernelGGL(k_copy2d<T>, dim3(gx, B), dim3(256), 0, s, n, n, (const T *)nullptr,
                           (int64_t)0, (int64_t)0, Mw, n, sM);
        MQR_LAUNCH_CHECK();
    }
    if (dQ)
        MQR_CHECK(bgemm(f64, true, false, (int)n, (int)n, (int)m, -1.0, dQ, lddq, sdq, Q, ldq, sq,
                        1.0, Mw, n, sM, B, K_GEMM, s));
    {
        ProfScope ps(K_OTHER, 0.0, 8.0 * batch * n * n, s);
        const int64_t nt = (n + CL_T - 1) / CL_T;
        hipLaunchKernelGGL(k_copyltu<T>, dim3((unsigned)(nt * (nt + 1) / 2), B), dim3(256), 0, s,
                           n, Mw, sM);
        hipLaunchKernelGGL(k_inv_diag_batched<T>, dim3(nblk, B), dim3(small::NT), 0, s, n, R, ldr,
                           sr, Iw);
        MQR_LAUNCH_CHECK();
    }
    // B = dQ + Q copyltu(M), in the workspace (m x n, ld m), dQ read as the
    // GEMM's beta term
    MQR_CHECK(bgemm(f64, false, false, (int)m, (int)n, (int)n, 1.0, Q, ldq, sq, Mw, n, sM,
                    dQ ? 1.0 : 0.0, Ww, m, sW, B, K_GEMM, s, dQ, lddq, sdq));
    // dA = B R^{-T}: blocks from the last
    for (int J = nblk - 1; J >= 0; --J) {
        const int64_t j0 = (int64_t)J * AB;
        const int bj = (int)std::min<int64_t>(AB, n - j0);
        // X_J = B_J R_JJ^{-T}, into dA
        MQR_CHECK(bgemm(f64, false, true, (int)m, bj, bj, 1.0, Ww + j0 * m, m, sW,
                        Iw + (int64_t)J * AB * AB, AB, sI, 0.0, dA + j0 * ldda, ldda, sda, B,
                        K_TRSM, s));
        // B_<J -= X_J R_<J,J^T
        if (j0 > 0)
            MQR_CHECK(bgemm(f64, false, true, (int)m, (int)j0, bj, -1.0, dA + j0 * ldda, ldda,
                            sda, R + j0 * ldr, ldr, sr, 1.0, Ww, m, sW, B, K_TRSM, s));
    }
    return MQR_SUCCESS;
}

}  // namespace

int qr_backward(mqr_handle_t hd, bool f64, int64_t batch, int64_t m, int64_t n, const void *Q,
                int64_t ldq, int64_t sq, const void *R, int64_t ldr, int64_t sr, const void *dQ,
                int64_t lddq, int64_t sdq, const void *dR, int64_t lddr, int64_t sdr, void *dA,
                int64_t ldda, int64_t sda) {
    if (f64)
        return qr_backward_t<double>(hd, batch, m, n, (const double *)Q, ldq, sq,
                                     (const double *)R, ldr, sr, (const double *)dQ, lddq, sdq,
                                     (const double *)dR, lddr, sdr, (double *)dA, ldda, sda);
    return qr_backward_t<float>(hd, batch, m, n, (const float *)Q, ldq, sq, (const float *)R, ldr,
                                sr, (const float *)dQ, lddq, sdq, (const float *)dR, lddr, sdr,
                                (float *)dA, ldda, sda);
}

}  // namespace mqr
