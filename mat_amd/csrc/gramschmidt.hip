// QR by Gram-Schmidt (QR(A, Q, R, QR_GramSchmidt), QR/QR.h:36-42 with
// QR/__GramSchmidt.h:2-18) on the GPU.
//
// The reference runs modified Gram-Schmidt column by column: n^2 / 2
// dependent dot products and axpys of length m, Q = A R^{-1} with R's
// diagonal positive.  That thin QR with a positive diagonal is unique, so
// the GPU forms the same Q and R by block classical Gram-Schmidt with
// reorthogonalisation (BCGS2) over blocks of nb columns, every projection a
// GEMM on the MFMA tiles:
//   S1 = Q_<^T W,  W -= Q_< S1,   W = Q~ R~  (panel QR, positive diagonal)
//   S2 = Q_<^T Q~, Q~ -= Q_< S2,  Q~ = Q_k R^ (panel QR, positive diagonal)
//   R(<, k) = S1 + S2 R~,  R(k, k) = R^ R~
// where Q_< holds the columns already orthonormalised and each block's own
// orthonormalisation is the Householder panel of the factorisation path,
// its Q formed explicitly and its column signs flipped so that R's diagonal
// is positive (Gram-Schmidt's R_jj = ||q_j|| > 0).  The second pass makes Q
// orthogonal to working precision; the reference's MGS loses orthogonality
// in proportion to cond(A), so the two agree to O(cond(A) u).
// Requires m >= n (the reference divides by zero norms beyond the rank).
#include <algorithm>

#include "mqr_internal.h"

namespace mqr {

int orgqr_f64(mqr_handle_t hd, int64_t m, int64_t nq, int64_t k, const double *VR, int64_t ldvr,
              const double *T, int64_t ldt, double *Q, int64_t ldq);

namespace {

// Rb (b x b, ld b) = D triu(VR[0:b, 0:b]) and d[j] = D_jj = sign(VR(j, j))
// (sign(0) = +1): the panel's R with its rows signed to a positive diagonal.
__global__ void k_gs_take_r(int b, const double *__restrict__ VR, int64_t ld,
                            double *__restrict__ Rb, double *__restrict__ d) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= b * b) return;
    const int r = e % b, c = e / b;
    const double sg = VR[(int64_t)r * ld + r] < 0.0 ? -1.0 : 1.0;
    Rb[e] = r <= c ? sg * VR[(int64_t)c * ld + r] : 0.0;
    if (c == 0) d[r] = sg;
}

// Q(:, j) *= d[j] for the m x b block Q (ld)
__global__ void k_gs_sign_cols(int64_t m, int b, double *__restrict__ Q, int64_t ld,
                               const double *__restrict__ d) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= m * b) return;
    const int64_t r = e % m;
    const int c = (int)(e / m);
    Q[c * ld + r] *= d[c];
}

struct GsWs {
    double *Pm, *S1, *S2, *Rt, *Rh, *d, *tau, *T, *P;
    int64_t pcap;
    void *pws;
};

// W (m x b, ld) = Q~ R~: Q~ overwrites W (orthonormal columns), R~ (b x b,
// ld b, positive diagonal) goes to Rout.
int block_orth(mqr_handle_t hd, int64_t m, int b, double *W, int64_t ld, double *Rout,
               const GsWs &w) {
    hipStream_t s = hd->stream;
    MQR_HIP(hipMemcpy2DAsync(w.Pm, sizeof(double) * m, W, sizeof(double) * ld, sizeof(double) * m,
                             b, hipMemcpyDeviceToDevice, s));
    MQR_CHECK(panel_factor(hd, m, b, w.Pm, m, w.tau, w.T, hd->nb, w.pws));
    hipLaunchKernelGGL(k_gs_take_r, dim3((unsigned)((b * b + 255) / 256)), dim3(256), 0, s, b,
                       (const double *)w.Pm, m, Rout, w.d);
    MQR_LAUNCH_CHECK();
    MQR_CHECK(orgqr_f64(hd, m, b, b, w.Pm, m, w.T, hd->nb, W, ld));
    hipLaunchKernelGGL(k_gs_sign_cols, dim3((unsigned)((m * b + 255) / 256)), dim3(256), 0, s, m,
                       b, W, ld, (const double *)w.d);
    MQR_LAUNCH_CHECK();
    return MQR_SUCCESS;
}

// S (j0 x b, ld j0) = Q(:, 0:j0)^T W, split-K through P
int project_coeffs(mqr_handle_t hd, int64_t m, int64_t j0, int b, const double *Q, int64_t ldq,
                   const double *W, double *S, const GsWs &w) {
    int splits = gemm_pick_splits((int)j0, b, (int)m, 4 * hd->cu_count);
    while (splits > 1 && (int64_t)splits * j0 * b > w.pcap) --splits;
    GemmF64 g;
    g.M = (int)j0; g.N = b; g.K = (int)m;
    g.ta = true;
    g.A = Q; g.lda = ldq; g.B = W; g.ldb = ldq;
    g.C = splits > 1 ? w.P : S; g.ldc = j0;
    g.splits = splits; g.split_stride = j0 * b; g.kind = K_VTA;
    MQR_CHECK(gemm_f64(g, hd->stream));
    if (splits > 1) MQR_CHECK(reduce_partials(j0 * b, w.P, splits, j0 * b, hd->stream, S));
    return MQR_SUCCESS;
}

// W -= Q(:, 0:j0) S
int project_out(mqr_handle_t hd, int64_t m, int64_t j0, int b, const double *Q, int64_t ldq,
                double *W, const double *S) {
    GemmF64 g;
    g.M = (int)m; g.N = b; g.K = (int)j0;
    g.A = Q; g.lda = ldq; g.B = S; g.ldb = j0;
    g.C = W; g.ldc = ldq; g.alpha = -1.0; g.beta = 1.0; g.kind = K_UPDATE;
    return gemm_f64(g, hd->stream);
}

}  // namespace

int qr_gramschmidt_f64(mqr_handle_t hd, int64_t m, int64_t n, const double *A, int64_t lda,
                       double *Q, int64_t ldq, double *R, int64_t ldr) {
    // blocks of the panel width; T is stored with the handle's block size
    // as leading dimension, the layout the Q formation reads
    const int nb = std::min(hd->nb, kMaxNb), nbt = hd->nb;
    hipStream_t s = hd->stream;
    const int64_t pcap = std::max<int64_t>((int64_t)64 * nb * n, 1);
    const size_t doubles = (size_t)(m * nb + 2 * n * nb + 2 * nb * nb + 2 * nb + (int64_t)nbt * nb +
                                    pcap) +
                           panel_cholqr_ws_bytes(m) / sizeof(double) + 8 * 64;
    MQR_CHECK(ws_reserve(hd, hd->ws_ir, sizeof(double) * doubles));
    GsWs w;
    double *p = (double *)hd->ws_ir.ptr;
    auto take = [&](int64_t cnt) {
        double *q = p;
        p = (double *)(((uintptr_t)(p + cnt) + 255) & ~(uintptr_t)255);
        return q;
    };
    w.Pm = take(m * nb);
    w.S1 = take(n * nb);
    w.S2 = take(n * nb);
    w.Rt = take((int64_t)nb * nb);
    w.Rh = take((int64_t)nb * nb);
    w.d = take(nb);
    w.tau = take(nb);
    w.T = take((int64_t)nbt * nb);
    w.P = take(pcap);
    w.pcap = pcap;
    w.pws = (void *)p;
    if (Q != A)  // in place (Q is A, QR/QR.h:39-40) needs no copy
        MQR_HIP(hipMemcpy2DAsync(Q, sizeof(double) * ldq, A, sizeof(double) * lda,
                                 sizeof(double) * m, n, hipMemcpyDeviceToDevice, s));
    MQR_CHECK(scale_f64(n, n, 0.0, R, ldr, s));
    for (int64_t j0 = 0; j0 < n; j0 += nb) {
        const int b = (int)std::min<int64_t>(nb, n - j0);
        double *W = Q + j0 * ldq;
        double *Rkk = R + j0 + j0 * ldr;
        if (j0 == 0) {
            MQR_CHECK(block_orth(hd, m, b, W, ldq, w.Rt, w));
            MQR_HIP(hipMemcpy2DAsync(Rkk, sizeof(double) * ldr, w.Rt, sizeof(double) * b,
                                     sizeof(double) * b, b, hipMemcpyDeviceToDevice, s));
            continue;
        }
        // first pass: project out the earlier columns, orthonormalise the block
        MQR_CHECK(project_coeffs(hd, m, j0, b, Q, ldq, W, w.S1, w));
        MQR_CHECK(project_out(hd, m, j0, b, Q, ldq, W, w.S1));
        MQR_CHECK(block_orth(hd, m, b, W, ldq, w.Rt, w));
        // second pass on the orthonormal block
        MQR_CHECK(project_coeffs(hd, m, j0, b, Q, ldq, W, w.S2, w));
        MQR_CHECK(project_out(hd, m, j0, b, Q, ldq, W, w.S2));
        MQR_CHECK(block_orth(hd, m, b, W, ldq, w.Rh, w));
        // R(0:j0, k) = S1 + S2 R~, R(k, k) = R^ R~
        GemmF64 g;
        g.M = (int)j0; g.N = b; g.K = b;
        g.A = w.S2; g.lda = j0; g.B = w.Rt; g.ldb = b;
        g.C = R + j0 * ldr; g.ldc = ldr;
        g.Cin = w.S1; g.ldcin = j0; g.alpha = 1.0; g.beta = 1.0; g.kind = K_W2;
        MQR_CHECK(gemm_f64(g, s));
        GemmF64 d;
        d.M = b; d.N = b; d.K = b;
        d.A = w.Rh; d.lda = b; d.B = w.Rt; d.ldb = b;
        d.C = Rkk; d.ldc = ldr; d.alpha = 1.0; d.beta = 0.0; d.kind = K_W2;
        MQR_CHECK(gemm_f64(d, s));
    }
    return MQR_SUCCESS;
}

}  // namespace mqr
