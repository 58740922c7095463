// FP64 blocked Householder QR drivers: factorisation, Q formation, Q / Q^T
// application, triangular solve and the plain least-squares solve.
//
// The factor loop follows QR/__Householder_Blocked.h:214-245: for each block
// of nb columns, the panel reflectors (:62-84), the T factor (:86-103) and the
// trailing update C <- (I - V T V^T)^T C (:143-212).  On the device the update
// is three launches: one MFMA GEMM computing V^T [V | C] in split-K partials
// (its first b columns are the Gram matrix V^T V the T recurrence needs),
// W2 = T^T (sum of partials), and the MFMA GEMM C -= V W2.
#include <algorithm>

#include "mqr_internal.h"

namespace mqr {
namespace {

__global__ void k_set_identity(int64_t m, int64_t n, double *Q, int64_t ldq) {
    const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= m * n) return;
    const int64_t r = idx % m, c = idx / m;
    Q[r + c * ldq] = r == c ? 1.0 : 0.0;
}

constexpr int TRB = 64;  // diagonal block of the triangular solve

// Solve the TRB x TRB (or smaller) diagonal block in place for up to 64
// right-hand sides per workgroup: upper (R x = b, backward) or its transpose
// (R^T x = b, forward).
__global__ __launch_bounds__(256) void k_trsm_diag(int nbk, int64_t nrhs, const double *R,
                                                   int64_t ldr, double *B, int64_t ldb,
                                                   int trans) {
    __shared__ double Rs[TRB][TRB + 1];
    __shared__ double Xs[TRB][65];
    const int tid = threadIdx.x;
    const int64_t c0 = (int64_t)blockIdx.x * 64;
    const int ncol = (int)std::min<int64_t>(64, nrhs - c0);
    for (int e = tid; e < nbk * nbk; e += 256) {
        const int r = e % nbk, c = e / nbk;
        Rs[r][c] = R[r + (int64_t)c * ldr];
    }
    for (int e = tid; e < nbk * ncol; e += 256) {
        const int r = e % nbk, c = e / nbk;
        Xs[r][c] = B[r + (c0 + c) * ldb];
    }
    __syncthreads();
    if (!trans) {
        for (int j = nbk - 1; j >= 0; --j) {
            if (tid < ncol) Xs[j][tid] = Xs[j][tid] / Rs[j][j];
            __syncthreads();
            for (int e = tid; e < j * ncol; e += 256) {
                const int i = e % j, c = e / j;
                Xs[i][c] = fma(-Rs[i][j], Xs[j][c], Xs[i][c]);
            }
            __syncthreads();
        }
    } else {
        for (int j = 0; j < nbk; ++j) {
            if (tid < ncol) Xs[j][tid] = Xs[j][tid] / Rs[j][j];
            __syncthreads();
            const int rest = nbk - j - 1;
            for (int e = tid; e < rest * ncol; e += 256) {
                const int i = j + 1 + e % rest, c = e / rest;
                Xs[i][c] = fma(-Rs[j][i], Xs[j][c], Xs[i][c]);
            }
            __syncthreads();
        }
    }
    for (int e = tid; e < nbk * ncol; e += 256) {
        const int r = e % nbk, c = e / nbk;
        B[r + (c0 + c) * ldb] = Xs[r][c];
    }
}

// One TRB-row block i0 of the back substitution R x = c (one right-hand
// side), with the diagonal blocks' inverses precomputed (Xd: the 64 x 64
// inverses side by side, ld 64).  Every workgroup forms x = Xd_q c[i0:i0+bs]
// (c already reduced by the later blocks) as a parallel product: four threads
// per row, each a quarter of the columns, summed in a fixed order, so no
// sequential elimination remains on the chain.  Workgroup 0 stores x to xout;
// workgroup w then subtracts R[r, i0:i0+bs] x from c[r] for its 64 rows
// r < i0.
__global__ __launch_bounds__(256) void k_trsv_inv_block(int64_t i0, int bs,
                                                        const double *__restrict__ R, int64_t ldr,
                                                        const double *__restrict__ Xd,
                                                        double *__restrict__ c,
                                                        double *__restrict__ xout) {
    __shared__ double cs[TRB], xs[TRB], part[4][TRB];
    const int tid = threadIdx.x, row = tid & (TRB - 1), qt = tid >> 6;
    const int64_t r = (int64_t)blockIdx.x * TRB + row;
    // every load first (c, the inverse's row piece and this thread's piece of
    // R's rows above the block), so they share one memory round trip
    double xv[16], rv[16];
    {
        const double *Xq = Xd + i0 * TRB;  // block q's inverse, ld TRB
#pragma unroll
        for (int u = 0; u < 16; ++u) {
            const int k = qt * 16 + u;
            xv[u] = (row < bs && k >= row && k < bs) ? Xq[row + (int64_t)k * TRB] : 0.0;
            rv[u] = (r < i0 && k < bs) ? R[r + (i0 + k) * ldr] : 0.0;
        }
    }
    if (tid < TRB) cs[tid] = tid < bs ? c[i0 + tid] : 0.0;
    __syncthreads();
    {
        double acc = 0.0;
        if (row < bs)
#pragma unroll
            for (int u = 0; u < 16; ++u) {
                const int k = qt * 16 + u;
                if (k >= row && k < bs) acc = fma(xv[u], cs[k], acc);
            }
        part[qt][row] = acc;
    }
    __syncthreads();
    if (tid < TRB) {
        const double xi = (part[0][tid] + part[1][tid]) + (part[2][tid] + part[3][tid]);
        xs[tid] = xi;
        if (blockIdx.x == 0 && tid < bs) xout[i0 + tid] = xi;
    }
    __syncthreads();
    // rows [64 b, 64 b + 64) above the block: four threads per row, each a
    // quarter of the block's columns, partial sums added in a fixed order
    double acc = 0.0;
    if (r < i0)
#pragma unroll
        for (int u = 0; u < 16; ++u) {
            const int j = qt * 16 + u;
            if (j < bs) acc = fma(rv[u], xs[j], acc);
        }
    part[qt][row] = acc;
    __syncthreads();
    if (tid < TRB && r < i0) c[r] -= (part[0][tid] + part[1][tid]) + (part[2][tid] + part[3][tid]);
}

__global__ void k_copy_vec(int64_t n, const double *__restrict__ x, double *__restrict__ y) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) y[i] = x[i];
}

int target_blocks(mqr_handle_t h) { return 4 * h->cu_count; }

// Y (h x b, ld h) = V op(T) on stream st, op(T) = T^T (trans) or T: the
// left factor of the block reflector, I - V T V^T = I - (V T) V^T.
int form_y(hipStream_t st, int64_t h, int b, const double *Vp, int64_t ldv, const double *Tp,
           int64_t ldt, bool trans, double *Y) {
    GemmF64 g;
    g.M = (int)h; g.N = b; g.K = b;
    g.tb = trans; g.amask = MASK_VLOWER;
    g.A = Vp; g.lda = ldv; g.B = Tp; g.ldb = ldt;
    g.C = Y; g.ldc = h; g.kind = K_W2;
    return gemm_f64(g, st);
}

// C <- C - Y (V^T C) for the h x nc matrix C on stream st, Y from form_y:
// applies (I - V T V^T)^T (Y = V T^T) or I - V T V^T (Y = V T) as in
// QR/__Householder_Blocked.h:143-212.  P (pcap doubles) receives the split-K
// partials of W = V^T C.
// W = V^T C (b x nc, ld b) into P[0 : b nc] on stream st; P (pcap doubles)
// holds the split-K partials.  fill: split K for four workgroups per CU
// (w_pick_splits; the factorisation's narrow updates), else as the other
// GEMMs (Q formation keeps the same splits, and so the same sums, for a thin
// and a full Q).
int block_w(mqr_handle_t hd, hipStream_t st, int64_t h, int b, const double *Vp, int64_t ldv,
            const double *C, int64_t ldc, int64_t nc, double *P, int64_t pcap, int kind_w,
            bool fill = false, int target = 0) {
    // W itself (splits = 1) must fit the partial buffer; callers chunk C
    if ((int64_t)b * nc > pcap) return MQR_ERR_INTERNAL;
    int splits = fill ? w_pick_splits(b, nc, h, hd->cu_count)
                      : gemm_pick_splits(b, (int)nc, (int)h, target > 0 ? target : target_blocks(hd));
    while (splits > 1 && (int64_t)splits * b * nc > pcap) --splits;
    GemmF64 g;
    g.M = b; g.N = (int)nc; g.K = (int)h;
    g.ta = true; g.amask = MASK_VLOWER;
    g.A = Vp; g.lda = ldv; g.B = C; g.ldb = ldc;
    g.C = P; g.ldc = b; g.alpha = 1.0; g.beta = 0.0;
    g.splits = splits; g.split_stride = (int64_t)b * nc; g.kind = kind_w;
    MQR_CHECK(gemm_f64(g, st));
    return reduce_partials((int64_t)b * nc, P, splits, (int64_t)b * nc, st);
}

int apply_block(mqr_handle_t hd, hipStream_t st, int64_t h, int b, const double *Vp, int64_t ldv,
                const double *Y, double *C, int64_t ldc, int64_t nc, double *P, int64_t pcap,
                int kind_w = K_VTA, int kind_u = K_UPDATE, bool fill = false, int target = 0) {
    if (nc <= 0) return MQR_SUCCESS;
    // column chunks whose W (b x chunk) fits the partial buffer P, so that a
    // right-hand side wider than the matrix (least squares with many
    // right-hand sides) never writes past P
    const int64_t chunk = std::max<int64_t>(1, pcap / b);
    for (int64_t c0 = 0; c0 < nc; c0 += chunk) {
        const int64_t ncc = std::min<int64_t>(chunk, nc - c0);
        double *Cc = C + c0 * ldc;
        MQR_CHECK(block_w(hd, st, h, b, Vp, ldv, Cc, ldc, ncc, P, pcap, kind_w, fill, target));
        GemmF64 u;
        u.M = (int)h; u.N = (int)ncc; u.K = b;
        u.A = Y; u.lda = h; u.B = P; u.ldb = b;
        u.C = Cc; u.ldc = ldc; u.alpha = -1.0; u.beta = 1.0; u.kind = kind_u;
        MQR_CHECK(gemm_f64(u, st));
    }
    return MQR_SUCCESS;
}

}  // namespace

namespace {
int merged_t(mqr_handle_t hd, hipStream_t st, int64_t h, int b, const double *V, int64_t ldv,
             const double *T, int64_t ldt, int64_t j0, double *G, double *Tb, double *tmp,
             double *P, int64_t pcap, int target = 0);
}  // namespace

int form_y_f64(hipStream_t st, int64_t h, int b, const double *Vp, int64_t ldv, const double *Tp,
               int64_t ldt, bool trans, double *Y) {
    return form_y(st, h, b, Vp, ldv, Tp, ldt, trans, Y);
}

int merged_t_f64(mqr_handle_t hd, hipStream_t st, int64_t h, int b, const double *V, int64_t ldv,
                 const double *T, int64_t ldt, int64_t j0, double *G, double *Tb, double *tmp,
                 double *P, int64_t pcap) {
    return merged_t(hd, st, h, b, V, ldv, T, ldt, j0, G, Tb, tmp, P, pcap);
}

// Workspace (doubles) the factorisation of an m x n matrix uses with panels
// of ib columns and outer blocks of ob: split-K partials of the near updates
// (side stream), of the deferred and merged updates (far stream) and of the
// next-block update, a ring of Y factors, the merged block reflector
// (G, T, its Y), the next block's W partials and the tall-panel scratch.
static int64_t w_parts_doubles(int64_t m) { return ((m + 63) / 64) * (int64_t)kMaxNb * kMaxNb; }

constexpr int kYRing = 8;  // Y factors kept alive: two outer blocks of four panels

struct GeqrfWs {
    double *P, *P2, *P3, *P4, *Wp, *G, *Tb, *tmp, *Yo;
    double *Y[kYRing];
    int64_t pcap, pcap2, pcap3, pcap4;
    void *pws;
    size_t doubles;
};

static GeqrfWs geqrf_carve(double *base, int64_t m, int64_t n, int ib, int ob) {
    GeqrfWs w{};
    double *p = base;
    auto take = [&](int64_t cnt) {
        double *q = p;
        p = (double *)(((uintptr_t)(p + cnt) + 255) & ~(uintptr_t)255);
        return q;
    };
    w.pcap = (int64_t)16 * ib * (n + ib);
    w.pcap2 = (int64_t)64 * ib * ib;
    w.P = take(w.pcap);
    w.P2 = take(w.pcap2);
    for (double *&y : w.Y) y = take(m * ib);
    w.Wp = take(2 * w_parts_doubles(m) + 4096 + 1024);  // [W | Gold] partials, C's top rows
    if (ob > ib) {
        w.pcap3 = std::max<int64_t>((int64_t)8 * ob * (n + ob), (int64_t)64 * ob * ob);
        w.pcap4 = (int64_t)64 * ib * (ob + ib);
        w.P3 = take(w.pcap3);
        w.P4 = take(w.pcap4);
        w.G = take((int64_t)ob * ob);
        w.Tb = take((int64_t)ob * ob);
        w.tmp = take((int64_t)ob * ob / 2 + ob);
        w.Yo = take(m * ob);
    }
    w.pws = (void *)p;
    w.doubles = (size_t)(p - base) + panel_cholqr_ws_bytes(m) / sizeof(double) + 64;
    return w;
}

int panel_factor(mqr_handle_t hd, int64_t h, int b, double *V, int64_t lda, double *tau,
                 double *T, int64_t ldt, void *pws) {
    if (h >= kCholqrMinRows) return panel_cholqr(hd, h, b, V, lda, tau, T, ldt, pws);
    return panel_householder(h, b, V, lda, tau, T, ldt, hd->stream);
}

int geqrf_f64(mqr_handle_t hd, int64_t m, int64_t n, double *A, int64_t lda, double *tau,
              double *T, int64_t ldt) {
    return geqrf_rhs_f64(hd, m, n, A, lda, tau, T, ldt, nullptr, 0, 0);
}

namespace {

// Events of one factorisation, destroyed on every return path.
struct FactorEvents {
    hipEvent_t far_done[kYRing] = {};  // side stream, after panel q's deferred update
    hipEvent_t merged_a[2] = {};       // far stream, outer block J's merged update of the
                                       // next-but-one block (slot J & 1)
    hipEvent_t ready = nullptr;        // far stream, end of its work
    ~FactorEvents() {
        for (hipEvent_t &x : far_done)
            if (x) (void)hipEventDestroy(x);
        for (hipEvent_t &x : merged_a)
            if (x) (void)hipEventDestroy(x);
        if (ready) (void)hipEventDestroy(ready);
    }
    int create() {
        for (hipEvent_t &x : far_done)
            MQR_HIP(hipEventCreateWithFlags(&x, hipEventDisableTiming));
        for (hipEvent_t &x : merged_a)
            MQR_HIP(hipEventCreateWithFlags(&x, hipEventDisableTiming));
        MQR_HIP(hipEventCreateWithFlags(&ready, hipEventDisableTiming));
        return MQR_SUCCESS;
    }
};

}  // namespace

// Blocked factorisation (QR/__Householder_Blocked.h:214-245) of A in place,
// applying every block reflector also to the m x nrhs matrix B (ldb), which
// ends as Q^T B.
//
// Panels of ib = min(nb, 64) columns, with a look-ahead of one panel: the
// handle's stream applies panel j to the next panel's columns and factors
// that panel (the critical chain) while other streams update the rest.
//
// Two-level blocking (nb = 64: outer blocks of 256 columns; nb = 128, 192,
// 256: outer blocks of nb): for panel j of outer block J (columns
// [c0, c1), the next block [c1, c2)):
//   side stream  applies panel j at K = ib to the rest of block J,
//   far stream   applies it to block J+1 (the deferred part), and after J's
//                last panel applies block J as ONE block reflector (T merged
//                from the panels' T blocks) to the far columns [c2, n) and to
//                B, at K = ob instead of ob / ib passes at K = ib.
// Far columns are thus read and written once per outer block.  The chain
// waits for the far stream only where its next panel starts a new outer
// block (that block's columns then need block J-1's merged update and block
// J's deferred parts), one outer block after the merged update was queued.
// With nb < 64 there is one level: the side stream applies each panel to
// every remaining column and to B.
// With nb > 64 the merged T of each outer block is what T returns (nb x nb
// blocks in the dgeqrt layout); otherwise T holds the panels' blocks.
int geqrf_rhs_f64(mqr_handle_t hd, int64_t m, int64_t n, double *A, int64_t lda, double *tau,
                  double *T, int64_t ldt, double *B, int64_t ldb, int64_t nrhs) {
    if (nrhs <= 0) B = nullptr;
    const int64_t k = std::min(m, n);
    if (k == 0) return MQR_SUCCESS;
    const int nbt = hd->nb;                              // T block width
    const int ib = std::min(nbt, kMaxNb);                // panel width
    // Outer blocks of 256 columns at the default block size pay where the far
    // columns are many: 8192^2 36.0 vs 37.3 ms, 16384x4096 least squares 24.3
    // vs 24.4 ms; at 32768 x 1024 the merges cost more than they save (5.4
    // vs 5.1 ms), so matrices 8 or more times taller than wide keep one level.
    const int ob = nbt > kMaxNb ? nbt
                                : (nbt == kMaxNb && m < 8 * n ? kDefaultOuter : ib);
    const bool two = ob > ib;
    GeqrfWs w = geqrf_carve(nullptr, m, n, ib, ob);
    MQR_CHECK(ws_reserve(hd, hd->ws, sizeof(double) * w.doubles));
    w = geqrf_carve((double *)hd->ws.ptr, m, n, ib, ob);
    FactorEvents ev;
    if (two) MQR_CHECK(ev.create());
    hipStream_t s0 = hd->stream, s1 = hd->side_stream, s2 = hd->far_stream;
    // the side and far streams run after everything queued on the handle's stream
    MQR_HIP(hipEventRecord(hd->ev_fork, s0));
    MQR_HIP(hipStreamWaitEvent(s1, hd->ev_fork, 0));
    if (two) MQR_HIP(hipStreamWaitEvent(s2, hd->ev_fork, 0));
    // outer block of column c: [c0, c1); the block after it ends at c2
    auto c1_of = [&](int64_t c) { return two ? std::min(n, (c / ob + 1) * ob) : n; };
    auto c2_of = [&](int64_t c) { return two ? std::min(n, (c / ob + 2) * ob) : n; };
    bool side_pending = false;  // the side stream's last update is not yet joined
    bool side_any = false;      // side-stream work was queued (joined at the end by ev_side)
    int64_t merged_upto = 0;    // outer blocks [0, merged_upto) have had their merged update
    int64_t s1_synced = 0;      // the side stream has waited for a merged update up to this block
    // the far stream's merged update of outer block J (all its panels
    // factored): T of the block, Y = V T^T, then [c2, n) and B.  The
    // next-but-one block's columns [c2, c2 + ob) go first, and
    // ev.merged_a[J & 1] marks them: the deferred parts of block J + 1's
    // panels (side stream) wait for those columns only, not for the rest.
    // the merged updates' split-K products aim at one workgroup per CU, not
    // four: they run beside the panel chain and the side stream, and fewer,
    // longer workgroups leave those more room (16384 x 4096 least squares
    // 19.27 -> 19.08 ms)
    const int far_target = hd->cu_count;
    auto merge_block = [&](int64_t J) -> int {
        const int64_t c0 = J * ob, bo = std::min<int64_t>(ob, k - c0);
        const int64_t h = m - c0;
        const int64_t c2 = std::min(n, c0 + 2 * ob);
        const double *V = A + c0 + c0 * lda;
        const bool far = c2 < n;
        if (!far && !B && nbt <= kMaxNb) {
            MQR_HIP(hipEventRecord(ev.merged_a[J & 1], s2));
            return MQR_SUCCESS;
        }
        const double *Tj = T + c0 * ldt;
        int64_t ldtj = ldt;
        if (bo > ib) {
            MQR_CHECK(merged_t(hd, s2, h, (int)bo, V, lda, T, ldt, c0, w.G, w.Tb, w.tmp, w.P3,
                               w.pcap3, far_target));
            Tj = w.Tb;
            ldtj = bo;
        }
        if (far || B) {
            MQR_CHECK(form_y(s2, h, (int)bo, V, lda, Tj, ldtj, true, w.Yo));
            if (far) {
                const int64_t na = std::min<int64_t>(ob, n - c2);
                MQR_CHECK(apply_block(hd, s2, h, (int)bo, V, lda, w.Yo, A + c0 + c2 * lda, lda,
                                      na, w.P3, w.pcap3, K_LOOKAHEAD, K_LOOKAHEAD, false,
                                      far_target));
                MQR_HIP(hipEventRecord(ev.merged_a[J & 1], s2));
                if (n - c2 > na)
                    MQR_CHECK(apply_block(hd, s2, h, (int)bo, V, lda, w.Yo,
                                          A + c0 + (c2 + na) * lda, lda, n - c2 - na, w.P3,
                                          w.pcap3, K_VTA, K_UPDATE, false, far_target));
            } else {
                MQR_HIP(hipEventRecord(ev.merged_a[J & 1], s2));
            }
            if (B)
                MQR_CHECK(apply_block(hd, s2, h, (int)bo, V, lda, w.Yo, B + c0, ldb, nrhs, w.P3,
                                      w.pcap3, K_APPLYQ, K_APPLYQ, false, far_target));
        }
        if (!(far || B)) MQR_HIP(hipEventRecord(ev.merged_a[J & 1], s2));
        // nb > 64: T returns the merged nb x nb blocks
        if (nbt > kMaxNb && bo > ib)
            MQR_HIP(hipMemcpy2DAsync(T + c0 * ldt, sizeof(double) * ldt, w.Tb,
                                     sizeof(double) * bo, sizeof(double) * bo, bo,
                                     hipMemcpyDeviceToDevice, s2));
        return MQR_SUCCESS;
    };
    // a tall full-width panel runs the fused CholeskyQR sequence, which also
    // forms Y = V T^T, takes the next-panel update into its first launch and
    // forms the following panel's W = V^T C in its last
    auto fused = [&](int64_t hp, int bp, int64_t bcols) {
        return hp >= kCholqrMinRows && bp == bcols;
    };
    auto fused_at = [&](int64_t c) {  // the panel starting at column c
        if (c >= k || c >= n) return false;
        return fused(m - c, (int)std::min<int64_t>(ib, k - c), std::min<int64_t>(ib, n - c));
    };
    // the panel at column c, after a panel of bprev columns, takes its first
    // Gram matrix from the previous panel's [W | Gold] (cholqr_fused_a)
    auto gold_at = [&](int64_t c, int64_t bprev) {
        return ib == kMaxNb && bprev == kMaxNb && fused_at(c) && k - c >= kMaxNb &&
               n - c >= kMaxNb && m - c <= kCholqrFusedMaxRows;
    };
    bool w_ready = false;  // Wp[0:4096] = V^T C of the current panel and the next one
    double *const ctop = w.Wp + 2 * w_parts_doubles(m);  // Cn's top rows (gold)
    bool g_ready = false;  // Wp[4096:8192] = that C's Gram matrix before the update
    int64_t b_next = 0;    // first panel column whose block is not yet applied to B (one level)
    const int b0 = (int)std::min<int64_t>(ib, k);
    if (fused(m, b0, std::min<int64_t>(ib, n))) {
        MQR_CHECK(cholqr_gram1(s0, m, b0, 0, nullptr, 0, nullptr, 0, A, lda, 0, w.pws));
        MQR_CHECK(cholqr_rest_a(s0, m, b0, A, lda, tau, T, ldt, w.pws));
        w_ready = b0 < n && fused_at(b0);
        g_ready = w_ready && gold_at(b0, b0);
        MQR_CHECK(cholqr_finish(s0, m, b0, A, lda, T, ldt, b0 < n ? w.Y[0] : nullptr,
                                A + b0 * lda, lda, (int)std::min<int64_t>(ib, n - b0),
                                w_ready ? w.Wp : nullptr, w.pws, false, nullptr, nullptr, nullptr,
                                0, g_ready ? ctop : nullptr));
    } else {
        MQR_CHECK(panel_factor(hd, m, b0, A, lda, tau, T, ldt, w.pws));
        if (b0 < n) MQR_CHECK(form_y(s0, m, b0, A, lda, T, ldt, true, w.Y[0]));
    }
    for (int64_t j = 0, it = 0; j < k; j += ib, ++it) {
        const int b = (int)std::min<int64_t>(ib, k - j);
        const int64_t h = m - j;
        const double *V = A + j + j * lda;
        const double *Y = w.Y[it % kYRing];
        const int64_t next = j + b;
        const int64_t c1 = c1_of(j), c2 = c2_of(j);
        const bool last_of_block = next >= c1 || next >= k;
        if (next >= n) {
            if (two && last_of_block) {
                MQR_HIP(hipEventRecord(hd->ev_fork, s0));
                MQR_HIP(hipStreamWaitEvent(s2, hd->ev_fork, 0));
                MQR_CHECK(merge_block(j / ob));
                merged_upto = j / ob + 1;
            }
            break;
        }
        const int64_t bn = std::min<int64_t>(ib, n - next);
        const int64_t rest0 = next + bn;
        const int64_t near_end = std::max(rest0, c1);  // side stream: [rest0, near_end)
        const int64_t def_end = std::max(near_end, c2); // far stream: [near_end, def_end)
        // panel j-1's side-stream update covers the next panel's columns
        if (side_pending) {
            MQR_HIP(hipStreamWaitEvent(s0, hd->ev_join, 0));
            side_pending = false;
        }
        // the side stream's deferred parts of the previous panels reach the
        // next panel's columns when the next panel starts a new outer block
        // (this block's deferred parts, after the previous block's merged
        // update of these columns), and when this panel is the first of its
        // block (the previous block's last deferred part)
        if (two && it > 0 && (j % ob == 0 || next >= c1))
            MQR_HIP(hipStreamWaitEvent(s0, ev.far_done[(it - 1) % kYRing], 0));
        const bool side_work = near_end > rest0 || (!two && B);
        const bool def_work = two && def_end > near_end;
        if (side_work || def_work || (two && last_of_block))
            MQR_HIP(hipEventRecord(hd->ev_fork, s0));
        if (side_work || def_work) {
            MQR_HIP(hipStreamWaitEvent(s1, hd->ev_fork, 0));
            // the panel after next (its first ib columns) first: the handle's
            // stream waits for those (ev_join) and not for the rest (ev_side)
            const int64_t first = std::min<int64_t>(ib, near_end - rest0);
            const bool split = near_end - rest0 > first;
            // (profiled as look-ahead: the bulk update is the roofline kernel)
            if (near_end > rest0)
                MQR_CHECK(apply_block(hd, s1, h, b, V, lda, Y, A + j + rest0 * lda, lda,
                                      split ? first : near_end - rest0, w.P, w.pcap,
                                      two ? K_NEAR : K_LOOKAHEAD, two ? K_NEAR : K_LOOKAHEAD,
                                      true));
            if (split) {
                MQR_HIP(hipEventRecord(hd->ev_join, s1));
                MQR_CHECK(apply_block(hd, s1, h, b, V, lda, Y, A + j + (rest0 + first) * lda,
                                      lda, near_end - rest0 - first, w.P, w.pcap,
                                      two ? K_NEAR : K_VTA, two ? K_NEAR : K_UPDATE, true));
            }
            if (!two && B) {
                MQR_CHECK(apply_block(hd, s1, h, b, V, lda, Y, B + j, ldb, nrhs, w.P, w.pcap,
                                      K_APPLYQ, K_APPLYQ));
                b_next = j + b;
            }
            if (!split) MQR_HIP(hipEventRecord(hd->ev_join, s1));
            // the deferred part: the next outer block's columns, which had
            // the previous block's merged update first (its first part,
            // ev.merged_a; the block before has no merged update)
            if (def_work) {
                const int64_t blk = j / ob;
                if (blk > 0 && s1_synced < blk) {
                    MQR_HIP(hipStreamWaitEvent(s1, ev.merged_a[(blk - 1) & 1], 0));
                    s1_synced = blk;
                }
                MQR_CHECK(apply_block(hd, s1, h, b, V, lda, Y, A + j + near_end * lda, lda,
                                      def_end - near_end, w.P4, w.pcap4, K_NEAR, K_NEAR, true));
            }
            MQR_HIP(hipEventRecord(hd->ev_side, s1));
            side_pending = true;
            side_any = true;
        }
        if (two) {
            // recorded for every panel: the side stream's work through this
            // panel's deferred part
            MQR_HIP(hipEventRecord(ev.far_done[it % kYRing], s1));
            if (last_of_block) {
                MQR_HIP(hipStreamWaitEvent(s2, hd->ev_fork, 0));
                MQR_CHECK(merge_block(j / ob));
                merged_upto = j / ob + 1;
            }
        }
        const int bnext = next < k ? (int)std::min<int64_t>(ib, k - next) : 0;
        if (next < k && fused(m - next, bnext, bn)) {
            double *Cn = A + j + next * lda;
            double *Xn = Cn + b;
            if (g_ready) {
                MQR_CHECK(cholqr_fused_a(s0, m - next, Y, h, w.Wp, ctop, Cn, lda, false, Xn,
                                         lda, tau + next, T + next * ldt, ldt, w.pws));
            } else {
                if (!w_ready)
                    MQR_CHECK(block_w(hd, s0, h, b, V, lda, Cn, lda, bn, w.P2, w.pcap2,
                                      K_LOOKAHEAD));
                MQR_CHECK(cholqr_gram1(s0, h, (int)bn, b, Y, h, w_ready ? w.Wp : w.P2,
                                       w_ready ? (int64_t)kMaxNb : (int64_t)b, Cn, lda, b,
                                       w.pws));
                MQR_CHECK(cholqr_rest_a(s0, m - next, bnext, Xn, lda, tau + next,
                                        T + next * ldt, ldt, w.pws));
            }
            // W of the panel after next needs this step's update of its
            // columns first: the side stream's (same outer block) or the far
            // stream's deferred part (next outer block)
            const int64_t next2 = next + bnext;
            w_ready = next2 < n && fused_at(next2);
            g_ready = w_ready && gold_at(next2, bnext);
            if (w_ready) {
                if (side_pending) {
                    MQR_HIP(hipStreamWaitEvent(s0, hd->ev_join, 0));
                    side_pending = false;
                }
                if (two && next2 >= c1) MQR_HIP(hipStreamWaitEvent(s0, ev.far_done[it % kYRing], 0));
            }
            MQR_CHECK(cholqr_finish(s0, m - next, bnext, Xn, lda, T + next * ldt, ldt,
                                    next2 < n ? w.Y[(it + 1) % kYRing] : nullptr,
                                    A + next + next2 * lda, lda,
                                    (int)std::min<int64_t>(ib, std::max<int64_t>(n - next2, 0)),
                                    w_ready ? w.Wp : nullptr, w.pws, false, nullptr, nullptr,
                                    nullptr, 0, g_ready ? ctop : nullptr));
            continue;
        }
        w_ready = false;
        g_ready = false;
        MQR_CHECK(apply_block(hd, s0, h, b, V, lda, Y, A + j + next * lda, lda, bn, w.P2, w.pcap2,
                              K_LOOKAHEAD, K_LOOKAHEAD));
        if (next < k) {
            double *Vn = A + next + next * lda;
            MQR_CHECK(panel_factor(hd, m - next, bnext, Vn, lda, tau + next, T + next * ldt, ldt,
                                   w.pws));
            if (next + bnext < n)
                MQR_CHECK(form_y(s0, m - next, bnext, Vn, lda, T + next * ldt, ldt, true,
                                 w.Y[(it + 1) % kYRing]));
        }
    }
    if (side_any) MQR_HIP(hipStreamWaitEvent(s0, hd->ev_side, 0));
    if (two) {
        // outer blocks whose last panel ended the loop early (m >= n: the
        // last panels have no columns to their right) still owe B and, with
        // nb > 64, their merged T
        const int64_t nblk = (k + ob - 1) / ob;
        if (merged_upto < nblk) {
            MQR_HIP(hipEventRecord(hd->ev_fork, s0));
            MQR_HIP(hipStreamWaitEvent(s2, hd->ev_fork, 0));
            for (int64_t J = merged_upto; J < nblk; ++J) MQR_CHECK(merge_block(J));
        }
        MQR_HIP(hipEventRecord(ev.ready, s2));
        MQR_HIP(hipStreamWaitEvent(s0, ev.ready, 0));
        return MQR_SUCCESS;
    }
    // one level: blocks whose step ended the loop early (the last one when m >= n)
    for (int64_t j = b_next; B && j < k; j += ib) {
        const int b = (int)std::min<int64_t>(ib, k - j);
        const double *V = A + j + j * lda;
        MQR_CHECK(form_y(s0, m - j, b, V, lda, T + j * ldt, ldt, true, w.Y[0]));
        MQR_CHECK(apply_block(hd, s0, m - j, b, V, lda, w.Y[0], B + j, ldb, nrhs, w.P, w.pcap,
                              K_APPLYQ, K_APPLYQ));
    }
    return MQR_SUCCESS;
}

int ormqr_f64(mqr_handle_t hd, bool trans, int64_t m, int64_t nrhs, int64_t k, const double *VR,
              int64_t ldvr, const double *T, int64_t ldt, double *C, int64_t ldc) {
    const int nb = hd->nb;
    const int64_t pcap = (int64_t)16 * nb * (nrhs + nb);
    MQR_CHECK(ws_reserve(hd, hd->ws, sizeof(double) * (pcap + (int64_t)m * nb + 1024)));
    double *P = (double *)hd->ws.ptr;
    double *Y = P + pcap;
    const int64_t nblk = (k + nb - 1) / nb;
    for (int64_t q = 0; q < nblk; ++q) {
        const int64_t j = trans ? q * nb : (nblk - 1 - q) * nb;
        const int b = (int)std::min<int64_t>(nb, k - j);
        const double *V = VR + j + j * ldvr;
        MQR_CHECK(form_y(hd->stream, m - j, b, V, ldvr, T + j * ldt, ldt, trans, Y));
        MQR_CHECK(apply_block(hd, hd->stream, m - j, b, V, ldvr, Y, C + j, ldc, nrhs, P, pcap,
                              K_APPLYQ, K_APPLYQ));
    }
    return MQR_SUCCESS;
}

namespace {

// Stored 64-column blocks applied together by orgqr_f64.
constexpr int kMergeBlocks = 4;

// Tb (nbig x nbig, ld nbig): the stored T blocks (T(0:b, j0 + c) for the
// block holding column j0 + c) on the diagonal 64-blocks, zero elsewhere.
__global__ void k_merged_t_init(int nbig, const double *__restrict__ T, int64_t ldt, int64_t j0,
                                double *__restrict__ Tb) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (int64_t)nbig * nbig) return;
    const int r = (int)(e % nbig), c = (int)(e / nbig);
    Tb[e] = ((r >> 6) == (c >> 6) && r <= c) ? T[(r & 63) + (j0 + c) * ldt] : 0.0;
}

// Tb (b x b, ld b): the compact-WY factor of the stored 64-column blocks
// j0, j0 + 64, ... (b columns in all) taken as one block reflector,
// I - V Tb V^T = H_j0 H_j0+64 ... with V the (h x b) unit-lower trapezoid.
// inv(Tb) = strict_upper(V^T V) + diag(1 / tau): its diagonal 64-blocks are
// the inverses of the stored T blocks and its off-diagonal blocks those of
// G = V^T V, so Tb's off-diagonal blocks follow by recursive doubling,
// Tb_12 = -T_1 G_12 T_2.  A reflector with tau = 0 has a zero row and column
// in its stored block, which the products carry into Tb.  G, tmp: b x b and
// b/2 x b doubles; P (pcap doubles) holds split-K partials.
int merged_t(mqr_handle_t hd, hipStream_t st, int64_t h, int b, const double *V, int64_t ldv,
             const double *T, int64_t ldt, int64_t j0, double *G, double *Tb, double *tmp,
             double *P, int64_t pcap, int target) {
    int splits = gemm_pick_splits(b, b, (int)h, target > 0 ? target : target_blocks(hd));
    while (splits > 1 && (int64_t)splits * b * b > pcap) --splits;
    GemmF64 g;
    g.M = b; g.N = b; g.K = (int)h;
    g.ta = true; g.amask = MASK_VLOWER; g.bmask = MASK_VLOWER; g.bmask_cols = b;
    g.A = V; g.lda = ldv; g.B = V; g.ldb = ldv;
    g.C = P; g.ldc = b; g.alpha = 1.0; g.beta = 0.0;
    g.splits = splits; g.split_stride = (int64_t)b * b; g.kind = K_APPLYQ;
    MQR_CHECK(gemm_f64(g, st));
    MQR_CHECK(reduce_partials((int64_t)b * b, P, splits, (int64_t)b * b, st, G));
    hipLaunchKernelGGL(k_merged_t_init, dim3((unsigned)(((int64_t)b * b + 255) / 256)), dim3(256),
                       0, st, b, T, ldt, j0, Tb);
    MQR_LAUNCH_CHECK();
    return upper_inverse_levels(hd, st, b, G, b, Tb, tmp);
}

// Events of one orgqr call, destroyed on every return path
struct LocalEvents {
    hipEvent_t e[4] = {};
    ~LocalEvents() {
        for (hipEvent_t &x : e)
            if (x) (void)hipEventDestroy(x);
    }
};

}  // namespace

// Q = H_0 H_1 ... applied to the first nq columns of I, from the last block
// to the first (QR/__Householder_Blocked.h:247-333), skipping the columns of
// Q that are still zero.  With 64-column blocks, four stored blocks are
// applied as one block reflector (merged_t), so that both products run with
// K = 256 instead of 64; the side stream forms the merged T and Y = V T of
// the next block while the handle's stream applies the current one (two
// buffers, ordered by events).
int orgqr_f64(mqr_handle_t hd, int64_t m, int64_t nq, int64_t k, const double *VR, int64_t ldvr,
              const double *T, int64_t ldt, double *Q, int64_t ldq) {
    const int nb = hd->nb;
    hipStream_t s0 = hd->stream, s1 = hd->side_stream;
    const int64_t total = m * nq;
    if (total > 0) {
        hipLaunchKernelGGL(k_set_identity, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s0,
                           m, nq, Q, ldq);
        MQR_LAUNCH_CHECK();
    }
    const int64_t kk = std::min(k, nq);
    if (nb != 64 || kk <= nb) {
        const int64_t pcap = (int64_t)16 * nb * (nq + nb);
        MQR_CHECK(ws_reserve(hd, hd->ws, sizeof(double) * (pcap + (int64_t)m * nb + 1024)));
        double *P = (double *)hd->ws.ptr;
        double *Y = P + pcap;
        const int64_t nblk = (kk + nb - 1) / nb;
        for (int64_t q = nblk - 1; q >= 0; --q) {
            const int64_t j = q * nb;
            const int b = (int)std::min<int64_t>(nb, kk - j);
            const double *V = VR + j + j * ldvr;
            // columns < j of Q[j:m, :] are still zero: skip them
            MQR_CHECK(form_y(s0, m - j, b, V, ldvr, T + j * ldt, ldt, false, Y));
            MQR_CHECK(apply_block(hd, s0, m - j, b, V, ldvr, Y, Q + j + j * ldq, ldq, nq - j, P,
                                  pcap, K_APPLYQ, K_APPLYQ));
        }
        return MQR_SUCCESS;
    }
    const int nbm = kMergeBlocks * nb;
    const int64_t sq = (int64_t)nbm * nbm;
    // split-K partials: a 256-row W = V^T Q over a tall, narrow Q needs up to
    // 64 splits to fill the GPU; the side stream's G = V^T V as many
    const int64_t pcap = std::max<int64_t>((int64_t)16 * nb * (nq + nb),
                                           std::min<int64_t>((int64_t)64 * nbm * nq, 1 << 24));
    const int64_t p2cap = 64 * sq;
    MQR_CHECK(ws_reserve(hd, hd->ws, sizeof(double) * (pcap + p2cap + 2 * (int64_t)m * nbm +
                                                       3 * sq + sq / 2 + 1024)));
    double *P = (double *)hd->ws.ptr;
    double *P2 = P + pcap;
    double *Ys[2] = {P2 + p2cap, P2 + p2cap + (int64_t)m * nbm};
    double *Tbs[2] = {Ys[1] + (int64_t)m * nbm, Ys[1] + (int64_t)m * nbm + sq};
    double *G = Tbs[1] + sq, *tmp = G + sq;
    LocalEvents ev;
    for (hipEvent_t &x : ev.e) MQR_HIP(hipEventCreateWithFlags(&x, hipEventDisableTiming));
    hipEvent_t *ready = ev.e, *freed = ev.e + 2;
    // the side stream runs after everything queued on the handle's stream
    MQR_HIP(hipEventRecord(hd->ev_fork, s0));
    MQR_HIP(hipStreamWaitEvent(s1, hd->ev_fork, 0));
    const int64_t nblk = (kk + nbm - 1) / nbm;
    // side stream: merged T and Y of block q into buffer q & 1, once the
    // update of block q + 2 has read that buffer
    auto prepare = [&](int64_t q) -> int {
        const int64_t j = q * nbm;
        const int b = (int)std::min<int64_t>(nbm, kk - j);
        const double *V = VR + j + j * ldvr;
        const int u = (int)(q & 1);
        if (q + 2 < nblk) MQR_HIP(hipStreamWaitEvent(s1, freed[u], 0));
        const double *Tj = T + j * ldt;
        int64_t ldtj = ldt;
        if (b > nb) {
            MQR_CHECK(merged_t(hd, s1, m - j, b, V, ldvr, T, ldt, j, G, Tbs[u], tmp, P2, p2cap));
            Tj = Tbs[u];
            ldtj = b;
        }
        MQR_CHECK(form_y(s1, m - j, b, V, ldvr, Tj, ldtj, false, Ys[u]));
        MQR_HIP(hipEventRecord(ready[u], s1));
        return MQR_SUCCESS;
    };
    MQR_CHECK(prepare(nblk - 1));
    for (int64_t q = nblk - 1; q >= 0; --q) {
        if (q > 0) MQR_CHECK(prepare(q - 1));
        const int64_t j = q * nbm;
        const int b = (int)std::min<int64_t>(nbm, kk - j);
        const int u = (int)(q & 1);
        MQR_HIP(hipStreamWaitEvent(s0, ready[u], 0));
        // columns < j of Q[j:m, :] are still zero: skip them
        MQR_CHECK(apply_block(hd, s0, m - j, b, VR + j + j * ldvr, ldvr, Ys[u], Q + j + j * ldq,
                              ldq, nq - j, P, pcap, K_APPLYQ, K_APPLYQ));
        MQR_HIP(hipEventRecord(freed[u], s0));
    }
    // the side stream's last work (block 0) was waited for above
    return MQR_SUCCESS;
}

// Back substitution R x = c for one right-hand side (c overwritten by x),
// one launch per TRB-row block from the last: every workgroup solves the
// diagonal block from c (final by then), then subtracts that block's columns
// times x from its share of the rows above; the block's x goes to a scratch
// vector (other workgroups of the launch still read c there) and is copied
// into c at the end.
int trsv_upper_f64(mqr_handle_t hd, int64_t n, const double *R, int64_t ldr, double *c) {
    // x (n + 64), then the diagonal blocks' inverses (64 x n, 256-byte aligned)
    const size_t xoff = ((size_t)(n + 64) * sizeof(double) + 255) & ~(size_t)255;
    MQR_CHECK(ws_reserve(hd, hd->ws_vec, xoff + sizeof(double) * (size_t)TRB * (size_t)n));
    double *x = (double *)hd->ws_vec.ptr;
    double *Xd = (double *)((char *)hd->ws_vec.ptr + xoff);
    const int64_t nblk = (n + TRB - 1) / TRB;
    {
        // all diagonal blocks inverted at once, in parallel
        ProfScope ps(K_TRSM, (double)n * TRB * TRB / 3.0, 16.0 * (double)n * TRB, hd->stream);
        MQR_CHECK(inv_diag_blocks_f64(hd->stream, n, R, ldr, Xd));
    }
    for (int64_t q = nblk - 1; q >= 0; --q) {
        const int64_t i0 = q * TRB;
        const int bs = (int)std::min<int64_t>(TRB, n - i0);
        const unsigned grid64 = (unsigned)std::max<int64_t>(1, (i0 + TRB - 1) / TRB);
        ProfScope ps(K_TRSM, (double)bs * bs + 2.0 * i0 * bs, 8.0 * ((double)bs * bs + i0 * bs),
                     hd->stream);
        hipLaunchKernelGGL(k_trsv_inv_block, dim3(grid64), dim3(256), 0, hd->stream, i0, bs, R, ldr,
                           (const double *)Xd, c, x);
        MQR_LAUNCH_CHECK();
    }
    hipLaunchKernelGGL(k_copy_vec, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, hd->stream, n,
                       (const double *)x, c);
    MQR_LAUNCH_CHECK();
    return MQR_SUCCESS;
}

int trsm_upper_f64(mqr_handle_t hd, bool trans, int64_t n, int64_t nrhs, const double *R,
                   int64_t ldr, double *B, int64_t ldb) {
    if (n <= 0 || nrhs <= 0) return MQR_SUCCESS;
    if (nrhs == 1 && !trans) return trsv_upper_f64(hd, n, R, ldr, B);
    const int64_t nblk = (n + TRB - 1) / TRB;
    const unsigned cgrid = (unsigned)((nrhs + 63) / 64);
    for (int64_t q = 0; q < nblk; ++q) {
        const int64_t i0 = trans ? q * TRB : (nblk - 1 - q) * TRB;
        const int bs = (int)std::min<int64_t>(TRB, n - i0);
        {
            ProfScope ps(K_TRSM, (double)bs * bs * nrhs, 8.0 * (bs * bs + 2.0 * bs * nrhs),
                         hd->stream);
            hipLaunchKernelGGL(k_trsm_diag, dim3(cgrid), dim3(256), 0, hd->stream, bs, nrhs,
                               R + i0 + i0 * ldr, ldr, B + i0, ldb, trans ? 1 : 0);
            MQR_LAUNCH_CHECK();
        }
        GemmF64 g;
        if (!trans) {
            if (i0 == 0) continue;
            // B[0:i0] -= R[0:i0, i0:i0+bs] X
            g.M = (int)i0; g.N = (int)nrhs; g.K = bs;
            g.A = R + i0 * ldr; g.lda = ldr; g.B = B + i0; g.ldb = ldb;
            g.C = B; g.ldc = ldb;
        } else {
            const int64_t rest = n - i0 - bs;
            if (rest <= 0) continue;
            // B[i0+bs:] -= R[i0:i0+bs, i0+bs:]^T X
            g.M = (int)rest; g.N = (int)nrhs; g.K = bs; g.ta = true;
            g.A = R + i0 + (i0 + bs) * ldr; g.lda = ldr; g.B = B + i0; g.ldb = ldb;
            g.C = B + i0 + bs; g.ldc = ldb;
        }
        g.alpha = -1.0; g.beta = 1.0; g.kind = K_TRSM;
        MQR_CHECK(gemm_f64(g, hd->stream));
    }
    return MQR_SUCCESS;
}

}  // namespace mqr
