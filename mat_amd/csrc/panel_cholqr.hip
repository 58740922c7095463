// Latency-optimal panel factorisation for tall panels.
//
// The reference factors each h x b panel column by column
// (QR/__Householder_Blocked.h:62-84): b dependent reductions over h rows.  On
// the GPU every such reduction is a grid-wide synchronisation, so the tall
// panel is factored instead by CholeskyQR2: two passes of
//   Gram (MFMA, split-K) -> Cholesky (+ triangular inverse) in one workgroup
//   -> Y = X R^{-1} (MFMA)
// (orthonormal to working accuracy when kappa(X) is below ~1e6).  The
// Householder vectors are then reconstructed from the orthonormal factor
// Q = Y R2^{-1} by an LU factorisation of Q(0:b, :) - S without pivoting,
// S = diag(+-1) chosen during the elimination (S_j = -sign of the current
// diagonal, sign(0) = +):  V1 = L, V2 = Q(b:h, :) U^{-1}, tau_j = -S_j U_jj,
// T = -U S L^{-T}, R = S R2 R1.  In exact arithmetic these are the
// reference's reflectors and R (R_jj = -sign(alpha_j) |R_jj|; a column with
// nothing left below the diagonal is reflected with tau = 2).
//
// A first-pass pivot d_i <= 1e-12 G_ii (a zero, dependent or nearly
// dependent column), a second-pass pivot below 1e-2, or a non-finite value
// raises a device flag; the rest of the sequence then leaves the panel
// untouched and the column-by-column Householder kernel (panel.hip) factors
// it instead, which also keeps the reference's tau = 0 for zero columns.
//
// The b x b factorisations run register-resident in one workgroup
// (small_dense.h); every barrier is executed by all of its waves.  The
// compact-WY factor T comes out of the reconstruction as well.
#include <climits>

#include "mqr_internal.h"
#include "small_dense.h"

namespace mqr {
namespace {

constexpr int PB = kMaxNb;  // padded panel width (small::N)
using SM = double[small::N + 1];

// Cholesky + inverse of the first pass.  G: reduced Gram (b x b, ld b).
// A pivot d_i <= 1e-12 G_ii (column i nearly in the span of the previous
// ones: kappa(X) beyond ~1e6, where two CholeskyQR passes no longer give an
// orthonormal factor to working accuracy), a zero column or a non-finite
// value sets the flag; the rest of the sequence then skips the panel.
__global__ __launch_bounds__(small::NT) void k_chol_inv(int b, const double *__restrict__ G,
                                                        double *__restrict__ R,
                                                        double *__restrict__ Rinv,
                                                        int *__restrict__ flag) {
    __shared__ small::Lds S[PB];
    __shared__ small::RowBuf rb;
    __shared__ double dv[2 * PB], thr[PB];
    __shared__ int fail;
    const int tid = threadIdx.x;
    if (tid == 0) fail = 0;
    small::lds_load(S, G, b, b);
    __syncthreads();
    if (tid < PB) thr[tid] = tid < b ? 1e-12 * S[tid][tid] : 0.0;
    __syncthreads();
    double a[small::RK];
    small::from_lds(a, S);
    small::chol(a, rb, dv, b, 0.0, &fail, thr);
    small::to_lds(a, S);
    __syncthreads();
    small::lds_store(S, R, PB, PB, PB);
    double x[small::RK];
    small::inv_upper(S, rb, dv, x);
    __syncthreads();
    small::to_lds(x, S);
    __syncthreads();
    small::lds_store(S, Rinv, PB, PB, PB);
    if (tid == 0) *flag = fail;
}

// Second pass and reconstruction.  Inputs: reduced Gram G2 of Z = X R1^{-1}
// (b x b, ld b), the top b rows of Z (ld ldz), R1 (padded, ld PB).  Outputs:
// the panel's top b x b block (R = S R2 R1 on/above the diagonal, V1 below),
// tau, the compact-WY factor T (ld ldt) and Minv = R2^{-1} U^{-1} (padded)
// for V2 = Z(b:h) Minv.
//
// With Q = Z R2^{-1}, Q1 = Q(0:b, :), and the LU factorisation
// Q1 - S = L U (S the signs), the reflectors I - V T V^T whose first b
// columns are Q S satisfy V1 = L, V2 = Q2 U^{-1} and T = -U S L^{-T}
// (top block: Q1 S - I = -V1 T V1^T).  T^T = L^{-1} (-S U^T) is one
// forward substitution.
__global__ __launch_bounds__(small::NT) void k_recon(int b, const double *__restrict__ G,
                                                     const double *__restrict__ Z, int64_t ldz,
                                                     const double *__restrict__ R1,
                                                     double *__restrict__ A, int64_t lda,
                                                     double *__restrict__ tau,
                                                     double *__restrict__ T, int64_t ldt,
                                                     double *__restrict__ Minv,
                                                     int *__restrict__ flag) {
    __shared__ small::Lds S0[PB], S1[PB], S2[PB], S3[PB];
    __shared__ small::RowBuf rb;
    __shared__ double sg[PB], dv[2 * PB];
    __shared__ int fail;
    if (*flag) return;
    const int tid = threadIdx.x, c = small::lane_id(), q = small::wave_id();
    if (tid == 0) fail = 0;
    small::lds_load(S0, Z, ldz, b);
    small::lds_load(S1, G, b, b);
    __syncthreads();
    double a[small::RK];
    // R2 = chol(G2) -> S2, R2^{-1} -> S1
    small::from_lds(a, S1);
    small::chol(a, rb, dv, b, 1e-2, &fail);
    small::to_lds(a, S2);
    __syncthreads();
    small::inv_upper(S2, rb, dv, a);
    __syncthreads();  // S1 (G2) no longer read
    small::to_lds(a, S1);
    __syncthreads();
    // Q1 = Z_top R2^{-1} -> S3
    small::matmul(S0, S1, false, true, [&](int r, int cc, double v) { S3[r][cc] = v; });
    __syncthreads();
    // LU of Q1 - S -> S3 (L strictly below, U on and above); U -> S0
    small::from_lds(a, S3);
    small::lu_sign(a, rb, dv, sg);
    small::to_lds(a, S3);
    small::keep_upper(a);
    small::to_lds(a, S0);
    __syncthreads();
    // U^{-1} -> S0
    small::inv_upper(S0, rb, dv, a);
    __syncthreads();
    small::to_lds(a, S0);
    __syncthreads();
    // Minv = R2^{-1} U^{-1} straight to global
    small::matmul(S1, S0, true, true, [&](int r, int cc, double v) { Minv[r + cc * PB] = v; });
    // T^T = L^{-1} (-S U^T): B(r, c) = -sg(r) U(c, r) for c <= r
#pragma unroll
    for (int j = 0; j < small::RK; ++j) {
        const int r = small::drow(j, q);
        a[j] = (c <= r) ? -sg[r] * S3[c][r] : 0.0;
    }
    small::solve_unit_lower(S3, rb, a);
    __syncthreads();  // S0 (U^{-1}) and S1 (R2^{-1}) no longer read
    small::to_lds(a, S0);
    small::lds_load(S1, R1, PB, PB);
    __syncthreads();
    for (int e = tid; e < PB * PB; e += small::NT) {
        const int r = e & (PB - 1), cc = e >> 6;
        if (r < b && cc < b) T[r + (int64_t)cc * ldt] = r <= cc ? S0[cc][r] : 0.0;
    }
    // R = S (R2 R1) with V1 (strictly below, from the LU) into the panel
    small::matmul(S2, S1, true, true, [&](int r, int cc, double v) {
        if (r < b && cc < b) A[r + (int64_t)cc * lda] = r <= cc ? sg[r] * v : S3[r][cc];
    });
    for (int i = tid; i < b; i += small::NT) tau[i] = -sg[i] * S3[i][i];
    __syncthreads();
    if (tid == 0 && fail) *flag = 1;
}

}  // namespace

// Workspace: Y (h x PB), R1, R1^{-1}, Minv (PB x PB), 64 Gram partials, flag.
size_t panel_cholqr_ws_bytes(int64_t h) {
    return sizeof(double) * ((size_t)h * PB + 3 * PB * PB + (size_t)64 * PB * PB) + 256;
}

// Factor the h x b panel X (ld) in place.
int panel_cholqr(mqr_handle_t hd, int64_t h, int b, double *X, int64_t ld, double *tau,
                 double *T, int64_t ldt, void *ws) {
    hipStream_t s = hd->stream;
    double *Y = (double *)ws;
    double *R1 = Y + (size_t)h * PB;
    double *I1 = R1 + PB * PB;
    double *Mi = I1 + PB * PB;
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
    // pass 1: Y = X R1^{-1}
    MQR_CHECK(gram(X, ld));
    {
        ProfScope ps(K_PANEL_SMALL, small_flops, small_bytes, s);
        hipLaunchKernelGGL(k_chol_inv, dim3(1), dim3(small::NT), 0, s, b, Gp, R1, I1, flag);
    }
    MQR_CHECK(apply(h, X, ld, I1, Y, h, flag));
    // pass 2 + reconstruction
    MQR_CHECK(gram(Y, h));
    {
        ProfScope ps(K_PANEL_SMALL, 4.0 * small_flops, small_bytes, s);
        hipLaunchKernelGGL(k_recon, dim3(1), dim3(small::NT), 0, s, b, Gp, Y, h, R1, X, ld, tau,
                           T, ldt, Mi, flag);
    }
    MQR_LAUNCH_CHECK();
    if (h > b) MQR_CHECK(apply(h - b, Y + b, h, Mi, X + b, ld, flag));
    // fallback: the column-by-column kernel runs only when the flag is up
    return panel_householder_gated(h, b, X, ld, tau, T, ldt, flag, s);
}

}  // namespace mqr
