// Latency-optimal panel factorisation for tall panels.
//
// The reference factors each h x b panel column by column
// (QR/__Householder_Blocked.h:62-84): b dependent reductions over h rows.  On
// the GPU every such reduction is a grid-wide synchronisation, so the tall
// panel is factored instead by shifted CholeskyQR3: three passes of
//   Gram (MFMA, split-K) -> Cholesky + triangular inverse (one wave)
//   -> Y = X R^{-1} (MFMA)
// with the first Gram shifted by 11 (h b + b (b + 1)) u ||X||_F^2.  The
// Householder vectors are then reconstructed from the orthonormal factor
// Q = Y R3^{-1} by an LU factorisation of Q(0:b, :) - S without pivoting,
// S = diag(+-1) chosen during the elimination (S_j = -sign of the current
// diagonal, sign(0) = +):  V1 = L, V2 = Q(b:h, :) U^{-1}, tau_j = -S_j U_jj,
// R = S R3 R2 R1.  In exact arithmetic these are the reference's reflectors
// and R (R_jj = -sign(alpha_j) |R_jj|; a column with nothing left below the
// diagonal is reflected with tau = 2).
//
// A Cholesky pivot at or below twice the shift (a zero or numerically
// dependent column) raises a device flag; the rest of the sequence then
// leaves the panel untouched and the column-by-column Householder kernel
// (panel.hip) factors it instead, which keeps the reference's tau = 0 for
// zero columns.
//
// The b x b factorisations run register-resident in one workgroup
// (small_dense.h); every barrier is executed by all of its waves.
#include <climits>

#include "mqr_internal.h"
#include "small_dense.h"

namespace mqr {
namespace {

constexpr int PB = kMaxNb;  // padded panel width (small::N)
using SM = double[small::N + 1];

// Cholesky + inverse of one pass (one wave).  G: reduced Gram (b x b, ld b).
// Pass 1 adds the shift and resets the flag; later passes are skipped once
// it is up.
__global__ __launch_bounds__(64) void k_chol_inv(int b, int64_t h, const double *__restrict__ G,
                                                 int pass, double *__restrict__ R,
                                                 double *__restrict__ Rinv,
                                                 int *__restrict__ flag) {
    __shared__ double rowbuf[PB], rdiag[PB];
    __shared__ SM S[PB];
    __shared__ int fail;
    if (pass > 1 && *flag) return;
    const int c = small::lane_id();
    if (c == 0) fail = 0;
    double shift = 0.0;
    if (pass == 1) {
        double tr = 0.0;
        for (int i = 0; i < b; ++i) tr += G[i + (int64_t)i * b];
        shift = 11.0 * ((double)h * b + (double)b * (b + 1)) * 1.1102230246251565e-16 * tr;
    }
    double a[small::N], x[small::N];
    small::load_col(a, G, b, b);
#pragma unroll
    for (int r = 0; r < small::N; ++r)
        if (r == c && c < b) a[r] += shift;
    __syncthreads();
    // pass 1: a pivot the shift dominates marks a dependent column; later
    // passes see nearly orthonormal columns
    small::chol_lds(a, rowbuf, b, pass == 1 ? 2.0 * shift : 1e-2, &fail, true);
    small::store_col(a, R, PB);
    small::col_to_lds(a, S);
    __syncthreads();
    rdiag[c] = 1.0 / S[c][c];
    __syncthreads();
    small::inv_upper_lds(S, rdiag, x);
    small::store_col(x, Rinv, PB);
    __syncthreads();
    if (c == 0) {
        if (pass == 1) *flag = fail;
        else if (fail) *flag = 1;
    }
}

// Final pass and reconstruction (256 threads; wave 0 runs the sequential
// factorisations, all waves the products).  Inputs: reduced Gram G3 of Y
// (b x b, ld b), the top b rows of Y (ld ldy), R1, R2 (padded, ld PB).
// Outputs: the panel's top b x b block (R on/above the diagonal, V1 below),
// tau, and Minv = R3^{-1} U^{-1} (padded) for V2 = Y(b:h) Minv.
__global__ __launch_bounds__(256) void k_recon(int b, const double *__restrict__ G,
                                               const double *__restrict__ Y, int64_t ldy,
                                               const double *__restrict__ R1,
                                               const double *__restrict__ R2,
                                               double *__restrict__ A, int64_t lda,
                                               double *__restrict__ tau,
                                               double *__restrict__ Minv,
                                               int *__restrict__ flag) {
    __shared__ SM S0[PB], S1[PB], S2[PB], S3[PB];
    __shared__ double sg[PB], rowbuf[PB], rdiag[PB];
    __shared__ int fail;
    if (*flag) return;
    const int c = small::lane_id(), w = small::wave_id(), tid = threadIdx.x;
    if (tid == 0) fail = 0;
    for (int e = tid; e < PB * PB; e += 256) {
        const int r = e % PB, cc = e / PB;
        S0[r][cc] = (r < b && cc < b) ? Y[r + (int64_t)cc * ldy] : (r == cc ? 1.0 : 0.0);
    }
    __syncthreads();
    // R3 = chol(G3) -> S2, R3^{-1} -> S1
    {
        double a[small::N], x[small::N];
        if (w == 0) small::load_col(a, G, b, b);
        small::chol_lds(a, rowbuf, b, 1e-2, &fail, w == 0);
        if (w == 0) small::col_to_lds(a, S2);
        __syncthreads();
        if (tid < PB) rdiag[tid] = 1.0 / S2[tid][tid];
        __syncthreads();
        if (w == 0) {
            small::inv_upper_lds(S2, rdiag, x);
            small::col_to_lds(x, S1);
        }
    }
    __syncthreads();
    // Q1 = Y_top R3^{-1} -> S3
    small::matmul_lds(S0, S1, S3, false);
    __syncthreads();
    // LU of Q1 - S -> S3 (L strictly below, U on and above); U -> S0
    if (w == 0) {
        double q[small::N];
        small::lds_to_col(q, S3);
        small::lu_sign(q, sg);
        small::col_to_lds(q, S3);
#pragma unroll
        for (int r = 0; r < small::N; ++r) q[r] = (r <= c) ? q[r] : 0.0;
        small::col_to_lds(q, S0);
    }
    __syncthreads();
    if (tid < PB) rdiag[tid] = 1.0 / S0[tid][tid];
    __syncthreads();
    // U^{-1} -> S0 (read U from S0 first, then overwrite)
    {
        double z[small::N];
        if (w == 0) small::inv_upper_lds(S0, rdiag, z);
        __syncthreads();
        if (w == 0) small::col_to_lds(z, S0);
    }
    __syncthreads();
    // Minv = R3^{-1} U^{-1} (upper) straight to global
#pragma unroll 1
    for (int q = 0; q < 16; ++q) {
        const int r = w * 16 + q;
        double acc = 0.0;
        for (int k = r; k <= c; ++k) acc = fma(S1[r][k], S0[k][c], acc);
        Minv[r + c * PB] = acc;
    }
    __syncthreads();
    // W = R3 R2 -> S1 (S0 <- R2)
    for (int e = tid; e < PB * PB; e += 256) S0[e % PB][e / PB] = R2[e];
    __syncthreads();
    small::matmul_lds(S2, S0, S1, true);
    __syncthreads();
    // R = S (W R1): write with V1 (strictly below, from the LU) into the panel
    for (int e = tid; e < PB * PB; e += 256) S0[e % PB][e / PB] = R1[e];
    __syncthreads();
#pragma unroll 1
    for (int q = 0; q < 16; ++q) {
        const int r = w * 16 + q;
        if (r >= b || c >= b) continue;
        if (r <= c) {
            double acc = 0.0;
            for (int k = r; k <= c; ++k) acc = fma(S1[r][k], S0[k][c], acc);
            A[r + (int64_t)c * lda] = sg[r] * acc;
        } else {
            A[r + (int64_t)c * lda] = S3[r][c];
        }
    }
    for (int i = tid; i < b; i += 256) tau[i] = -sg[i] * S3[i][i];
    __syncthreads();
    if (tid == 0 && fail) *flag = 1;
}

}  // namespace

// Workspace: Y, Z (h x PB each), R1, R2, R1^{-1}, R2^{-1}, Minv (PB x PB),
// 64 Gram partials, flag.
size_t panel_cholqr_ws_bytes(int64_t h) {
    return sizeof(double) * ((size_t)2 * h * PB + 5 * PB * PB + (size_t)64 * PB * PB) + 256;
}

// Factor the h x b panel X (ld) in place.
int panel_cholqr(mqr_handle_t hd, int64_t h, int b, double *X, int64_t ld, double *tau,
                 void *ws) {
    hipStream_t s = hd->stream;
    double *Y = (double *)ws;
    double *Z = Y + (size_t)h * PB;
    double *R1 = Z + (size_t)h * PB;
    double *R2 = R1 + PB * PB;
    double *I1 = R2 + PB * PB;
    double *I2 = I1 + PB * PB;
    double *Mi = I2 + PB * PB;
    double *Gp = Mi + PB * PB;
    int *flag = (int *)(Gp + (size_t)64 * PB * PB);

    int splits = gemm_pick_splits(b, b, (int)h, 4 * hd->cu_count);
    if (splits > 64) splits = 64;
    // Gram of W reduced into Gp[0 : b*b]
    auto gram = [&](const double *W, int64_t ldw) -> int {
        GemmF64 g;
        g.M = b; g.N = b; g.K = (int)h; g.ta = true;
        g.A = W; g.lda = ldw; g.B = W; g.ldb = ldw;
        g.C = Gp; g.ldc = b; g.splits = splits; g.split_stride = (int64_t)b * b;
        g.kind = K_PANEL;
        MQR_CHECK(gemm_f64(g, s));
        return reduce_partials((int64_t)b * b, Gp, splits, (int64_t)b * b, s);
    };
    // out(rows x b) = in(rows x b) * Tinv (padded, ld PB)
    auto apply = [&](int64_t rows, const double *in, int64_t ldi, const double *Tinv, double *out,
                     int64_t ldo, const int *skip) -> int {
        GemmF64 g;
        g.M = (int)rows; g.N = b; g.K = b;
        g.A = in; g.lda = ldi; g.B = Tinv; g.ldb = PB;
        g.C = out; g.ldc = ldo; g.kind = K_PANEL; g.skip_if = skip;
        return gemm_f64(g, s);
    };
    const double bb = b;
    const double small_flops = bb * bb * bb / 3.0, small_bytes = 8.0 * bb * bb * 3.0;
    // pass 1 (shifted)
    MQR_CHECK(gram(X, ld));
    {
        ProfScope ps(K_PANEL, small_flops, small_bytes, s);
        hipLaunchKernelGGL(k_chol_inv, dim3(1), dim3(64), 0, s, b, h, Gp, 1, R1, I1, flag);
    }
    MQR_CHECK(apply(h, X, ld, I1, Y, h, nullptr));
    // pass 2
    MQR_CHECK(gram(Y, h));
    {
        ProfScope ps(K_PANEL, small_flops, small_bytes, s);
        hipLaunchKernelGGL(k_chol_inv, dim3(1), dim3(64), 0, s, b, h, Gp, 2, R2, I2, flag);
    }
    MQR_CHECK(apply(h, Y, h, I2, Z, h, nullptr));
    // pass 3 + reconstruction
    MQR_CHECK(gram(Z, h));
    {
        ProfScope ps(K_PANEL, 4.0 * small_flops, small_bytes, s);
        hipLaunchKernelGGL(k_recon, dim3(1), dim3(256), 0, s, b, Gp, Z, h, R1, R2, X, ld, tau,
                           Mi, flag);
    }
    MQR_LAUNCH_CHECK();
    if (h > b) MQR_CHECK(apply(h - b, Z + b, h, Mi, X + b, ld, flag));
    // fallback: the column-by-column kernel runs only when the flag is up
    return panel_householder_gated(h, b, X, ld, tau, flag, s);
}

}  // namespace mqr
