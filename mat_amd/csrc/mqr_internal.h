// Internal declarations shared by the HIP translation units of libmatqr.
#pragma once

#include <rocprofiler-sdk-roctx/roctx.h>
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

#include "../../include/mat_qr.h"

namespace mqr {

// Status helpers --------------------------------------------------------------
void set_last_hip_error(hipError_t e, const char *where);

#define MQR_HIP(call)                                        \
    do {                                                     \
        hipError_t e_ = (call);                              \
        if (e_ != hipSuccess) {                              \
            ::mqr::set_last_hip_error(e_, #call);            \
            return MQR_ERR_HIP;                              \
        }                                                    \
    } while (0)

#define MQR_CHECK(expr)                 \
    do {                                \
        int s_ = (expr);                \
        if (s_ != MQR_SUCCESS) return s_; \
    } while (0)

#define MQR_LAUNCH_CHECK() MQR_HIP(hipGetLastError())

// Device workspace owned by a handle: grows, never shrinks until destroy.
struct Workspace {
    void *ptr = nullptr;
    size_t bytes = 0;
};

// Per-kernel-kind timing with HIP event pairs (enabled per handle).
enum KernelKind {
    K_PANEL = 0,        // GEMMs and reductions inside the panel factorisation
    K_VTA = 1,          // split-K W = V^T C partials
    K_PANEL_SMALL = 2,  // single-workgroup panel kernels (b x b factors, column panel)
    K_W2 = 3,         // W2 = op(T) W
    K_UPDATE = 4,     // C -= V W2 (trailing update / Q formation)
    K_TRSM = 5,       // triangular solves
    K_GEMM = 6,       // plain GEMM calls
    K_RESID = 7,      // refinement residual passes
    K_CAST = 8,       // precision conversions
    K_OTHER = 9,
    K_LOOKAHEAD = 10,   // next-block update on the critical stream
    K_APPLYQ = 11,      // Q / Q^T applications (ormqr, orgqr)
    K_NEAR = 12,        // one panel's update of the columns up to the next outer block's end
    K_COUNT = 13
};
const char *kernel_kind_name(int k);

struct Profiler {
    bool on = false;
    unsigned mask = ~0u;  // kernel kinds timed (bit k = KernelKind k)
    struct Rec {
        int kind;
        double flops, bytes;
        hipEvent_t a, b;
    };
    std::vector<hipEvent_t> pool;
    size_t used = 0;
    std::vector<Rec> recs;
};

// Active profiler of the calling thread (null when profiling is off).
extern thread_local Profiler *t_prof;

// Records an event pair around the launches made while it is alive.
struct ProfScope {
    Profiler *p;
    size_t idx;
    hipStream_t s;
    ProfScope(int kind, double flops, double bytes, hipStream_t stream);
    ~ProfScope();
};

// A named roctx range over the host-side span of an entry point or phase
// (`rocprofv3 --marker-trace` shows it beside the kernels it enqueued).
struct Range {
    explicit Range(const char *name) { roctxRangePushA(name); }
    ~Range() { roctxRangePop(); }
    Range(const Range &) = delete;
    Range &operator=(const Range &) = delete;
};

}  // namespace mqr

struct mqr_handle_st {
    int device = 0;
    hipStream_t own_stream = nullptr;
    // every kernel runs on `stream` (= own_stream, non-blocking); work is
    // ordered after and before the caller's user_stream with ev_in / ev_out
    hipStream_t stream = nullptr;
    hipStream_t user_stream = nullptr;
    hipEvent_t ev_in = nullptr, ev_out = nullptr;
    // look-ahead: trailing updates run on side_stream, ordered against
    // `stream` with ev_fork / ev_join; every call joins before returning.
    // A side update that starts with the columns the handle's stream reads
    // next records ev_join after those and ev_side at its end.
    hipStream_t side_stream = nullptr;
    hipEvent_t ev_fork = nullptr, ev_join = nullptr, ev_side = nullptr;
    // the refinement's host reads wait on ev_host (recorded after the copy of
    // the scalars) while later work is already queued behind it
    hipEvent_t ev_host = nullptr;
    // two-level blocking: an outer block's merged update of the far columns
    // runs on far_stream (also default priority)
    hipStream_t far_stream = nullptr;
    int nb = 64;                   // factorisation block size
    // block size each recent mqr_dgeqrf wrote its T blocks with (keyed by
    // the T pointer), so that a Q application after mqr_set_block_size is
    // rejected instead of reading T with the wrong block layout
    struct TLayout {
        const void *T = nullptr;
        int nb = 0;
    } t_layout[16];
    int t_layout_next = 0;
    int cu_count = 256;
    mqr::Workspace ws;             // general scratch
    mqr::Workspace ws_ir;          // mixed-precision / refinement scratch
    mqr::Workspace ws_vec;         // vector scratch of the triangular solve
    mqr::Workspace ws_lu;          // pivot / info scratch of the LU path
    double *host_pinned = nullptr; // small pinned buffer for device->host scalars
    mqr::Profiler prof;
};

namespace mqr {

// Ensure ws holds at least `bytes`; may synchronise the stream.
int ws_reserve(mqr_handle_t h, Workspace &ws, size_t bytes);

// Tile coordinates of the calling workgroup in an XCD-aware order.  The
// dispatcher deals consecutive workgroup ids round-robin to the eight XCDs
// (id % 8), each with its own L2.  Renumbering so that XCD x runs the
// contiguous range [x per, (x + 1) per) of the x-fastest tile order puts
// neighbouring tiles, which share operand slices (the row tiles of
// W = V^T C read the same slice of C), on one XCD and close in time.  Ids
// past 8 per keep their number.  A bijection of the grid: every tile is
// still computed exactly once, with the same arithmetic.
__device__ __forceinline__ void xcd_tile(int &bx, int &by, int &bz) {
    const unsigned gx = gridDim.x, gy = gridDim.y, gz = gridDim.z;
    const unsigned per = gx * gy * gz / 8;
    unsigned id = blockIdx.x + gx * (blockIdx.y + gy * blockIdx.z);
    if (id < per * 8) id = (id % 8) * per + id / 8;
    bx = (int)(id % gx);
    const unsigned t = id / gx;
    by = (int)(t % gy);
    bz = (int)(t / gy);
}

// ---- FP64 MFMA GEMM -------------------------------------------------------
// C[M x N] = alpha * op(A) * op(B) + beta * C, column-major.
// Mask modes apply to the *stored* operand before op():
//   MASK_NONE, MASK_VLOWER (unit lower trapezoid: zero above the diagonal,
//   one on it) for stored columns < mask_cols.
enum { MASK_NONE = 0, MASK_VLOWER = 1 };

struct GemmF64 {
    int M = 0, N = 0, K = 0;
    bool ta = false, tb = false;
    int amask = MASK_NONE, bmask = MASK_NONE;
    int bmask_cols = 0;  // B mask applies to stored columns < bmask_cols
    // op(A) upper triangular (op(A)(i, k) = 0 for k < i) / op(B) upper
    // triangular (op(B)(k, j) = 0 for k > j): each tile skips the K range
    // whose products are zero
    bool a_upper = false, b_upper = false;
    // 64 x 64 output tiles even when M, N >= 128 (more workgroups for small
    // grids)
    bool small_tiles = false;
    const double *A = nullptr;
    int64_t lda = 0;
    const double *B = nullptr;
    int64_t ldb = 0;
    double *C = nullptr;
    int64_t ldc = 0;
    // when set, beta C is read from Cin (ld ldcin, batch stride strideCin)
    // instead of C: C = alpha op(A) op(B) + beta Cin
    const double *Cin = nullptr;
    int64_t ldcin = 0, strideCin = 0;
    double alpha = 1.0, beta = 0.0;
    // split-K: when splits > 1, partial products are written (not accumulated)
    // to C + s * split_stride for split s; beta is ignored.
    int splits = 1;
    int64_t split_stride = 0;
    int kind = K_GEMM;  // profiler bucket
    // when non-null and *skip_if != 0 at run time, the launch does nothing
    const int *skip_if = nullptr;
    // strided batch: `batch` products, operand i at A + i * strideA, ...
    int batch = 1;
    int64_t strideA = 0, strideB = 0, strideC = 0;
};
int gemm_f64(const GemmF64 &g, hipStream_t s);

// FP32 MFMA GEMM (no masks): same conventions as GemmF64.
struct GemmF32 {
    int M = 0, N = 0, K = 0;
    bool ta = false, tb = false;
    const float *A = nullptr;
    int64_t lda = 0;
    const float *B = nullptr;
    int64_t ldb = 0;
    float *C = nullptr;
    int64_t ldc = 0;
    // when set, beta C is read from Cin (ld ldcin, batch stride strideCin)
    const float *Cin = nullptr;
    int64_t ldcin = 0, strideCin = 0;
    float alpha = 1.f, beta = 0.f;
    int splits = 1;
    int64_t split_stride = 0;
    int kind = K_GEMM;
    int batch = 1;   // strided batch as in GemmF64
    int64_t strideA = 0, strideB = 0, strideC = 0;
    // 0: FP32 MFMA; 1: BF16 products (hi * hi); 3: split BF16 (hi * hi +
    // hi * lo + lo * hi); always FP32 accumulation
    int bf16_terms = 0;
};
int gemm_f32(const GemmF32 &g, hipStream_t s);
// C = beta * C (beta == 0 writes zeros).
int scale_f64(int64_t m, int64_t n, double beta, double *C, int64_t ldc, hipStream_t s);
// Number of K splits giving roughly `target_blocks` workgroups.
int gemm_pick_splits(int M, int N, int K, int target_blocks);
// K splits of W = V^T C (b x nc, K = h rows): enough workgroups for four per
// CU (of `cus`), K chunks of at least min_rows rows, at most 256 splits.  A
// narrow C (the next panel's 64 columns) is one tile column, which
// gemm_pick_splits' 256-row chunks would leave on a quarter of the CUs.
int w_pick_splits(int b, int64_t nc, int64_t h, int cus, int min_rows = 128);

// ---- panel / T / block reflector ------------------------------------------
// Every panel routine factors the h x b panel P (ld) in place (R on and above
// the diagonal, reflectors below), and writes tau[0:b] and the b x b
// compact-WY factor T (ld ldt, zero below the diagonal).
// Householder panel, single workgroup.
int panel_householder(int64_t h, int b, double *P, int64_t ld, double *tau, double *T,
                      int64_t ldt, hipStream_t s);
// CholeskyQR2 + Householder reconstruction panel; on breakdown the
// reconstruction kernel factors the panel column by column itself.
// ws: panel_cholqr_ws_bytes(h).
size_t panel_cholqr_ws_bytes(int64_t h);
int panel_cholqr(mqr_handle_t hd, int64_t h, int b, double *X, int64_t ld, double *tau,
                 double *T, int64_t ldt, void *ws);
// The same in pieces, for callers that fuse the next-block update into the
// first and the next W = V^T C into the last: cholqr_gram1 applies
// C (hc x bn) -= Y W (kin > 0; Y ld ldy, W kin x bn, ld ldw), forms the Gram
// matrix of rows [g0, hc) of C (the panel) and its first Cholesky pass;
// cholqr_rest_a factors that panel (h = hc - g0 rows at X = C + g0) up to V1,
// R, tau and T; cholqr_finish writes V2 and, when Yf, Yf (h x b, ld h) =
// V T^T and, when Wp, W = V^T Cn (b x bn, ld 64) into Wp[0 : 4096] (Wp holds
// ceil(h / 64) x 4096 doubles).  cholqr_rest = cholqr_rest_a + cholqr_finish
// without W.
// C is FP64, or FP32 when c_f32 (then Xo, ld ldxo, receives the panel rows
// in FP64 for the factorisation).
int cholqr_gram1(hipStream_t s, int64_t hc, int bn, int kin, const double *Y, int64_t ldy,
                 const double *W, int64_t ldw, void *C, int64_t ldc, int64_t g0, void *ws,
                 bool c_f32 = false, double *Xo = nullptr, int64_t ldxo = 0);
int cholqr_rest_a(hipStream_t s, int64_t h, int b, double *X, int64_t ld, double *tau, double *T,
                  int64_t ldt, void *ws);
// Cn is FP64, or FP32 when cn_f32; with Yf, V32 and Y32 (ld h) receive FP32
// copies of V (unit-lower V1 explicit) and Yf, and Rout (ldr) the panel's
// b x b R block (zero below the diagonal).
// With ctop (and Wp), Wp[4096 : 8192] also receives Cn^T Cn, the Gram of
// the next block's columns before this panel's update (Wp then holds
// 2 x ceil(h / 64) x 4096 doubles), and ctop (64 x 64, ld 64, outside Wp)
// Cn's top 64 rows, for cholqr_fused_a.
int cholqr_finish(hipStream_t s, int64_t h, int b, double *X, int64_t ld, const double *T,
                  int64_t ldt, double *Yf, const void *Cn, int64_t ldcn, int bn, double *Wp,
                  void *ws, bool cn_f32 = false, float *V32 = nullptr, float *Y32 = nullptr,
                  double *Rout = nullptr, int64_t ldr = 0, double *ctop = nullptr);
// cholqr_gram1 + cholqr_rest_a for a 64-column panel of h <= 16384 rows after
// a 64-column panel, when that panel's cholqr_finish formed [W | Gold] (WG):
// the update and the first pass in one launch, the first Gram matrix from
// Gold - R12^T R12 (C: the panel's columns from the previous panel's top
// row, hc = h + 64 rows; X = the panel, C + 64 for FP64 C, its FP64 copy
// for FP32 C).
// one_pass: the caller tolerates an orthonormal factor accurate to about
// 1e-9 (the mixed-precision factorisation, whose FP32 updates perturb it by
// more); the second CholeskyQR pass is then skipped, on the device, for
// panels whose first factor predicts that accuracy.
int cholqr_fused_a(hipStream_t s, int64_t h, const double *Y, int64_t ldy, const double *WG,
                   const double *Ctop, void *C, int64_t ldc, bool c_f32, double *X, int64_t ldx,
                   double *tau, double *T, int64_t ldt, void *ws, bool one_pass = false);
int cholqr_rest(hipStream_t s, int64_t h, int b, double *X, int64_t ld, double *tau, double *T,
                int64_t ldt, double *Yf, void *ws);
// Blocked FP64 QR of A (m x n, lda) in place with tau and the T blocks (ldt);
// the block reflectors are also applied to B (m x nrhs, ldb), which ends as
// Q^T B (B may be null).
int geqrf_rhs_f64(mqr_handle_t hd, int64_t m, int64_t n, double *A, int64_t lda, double *tau,
                  double *T, int64_t ldt, double *B, int64_t ldb, int64_t nrhs);
// Panels at least this tall use panel_cholqr (a few launches, ~0.2 ms); the
// column-by-column kernel takes ~1.4 ms on a 64-wide panel of a few hundred
// rows (two barriers and a global round trip per 8-column round), so it is
// kept for panels under one 64-row chunk and for breakdown.
constexpr int64_t kCholqrMinRows = 64;
// Tallest panel of cholqr_fused_a (one 64-row chunk per workgroup, at most
// 256 workgroups).
constexpr int64_t kCholqrFusedMaxRows = 256 * 64;
// Either of the above by height.
int panel_factor(mqr_handle_t hd, int64_t h, int b, double *V, int64_t lda, double *tau,
                 double *T, int64_t ldt, void *ws);
// Y (h x b, ld h) = V op(T) (op = transpose when trans): the left factor of
// the block reflector I - V T V^T = I - (V T) V^T.
int form_y_f64(hipStream_t st, int64_t h, int b, const double *Vp, int64_t ldv, const double *Tp,
               int64_t ldt, bool trans, double *Y);
// Tb (b x b, ld b): the compact-WY factor of the stored 64-column panels
// j0, j0 + 64, ... (b columns in all, T blocks T(0:64, j0 + c), ld ldt) taken
// as one block reflector, from G = V^T V (split-K partials in P, pcap
// doubles) by recursive doubling; G: b x b, tmp: b/2 x b + b doubles.
int inv_diag_blocks_f64(hipStream_t s, int64_t n, const double *R, int64_t ldr, double *Xd);
int merged_t_f64(mqr_handle_t hd, hipStream_t st, int64_t h, int b, const double *V, int64_t ldv,
                 const double *T, int64_t ldt, int64_t j0, double *G, double *Tb, double *tmp,
                 double *P, int64_t pcap);
// The off-diagonal blocks of X = M^{-1} (ld n) for an upper n x n M whose
// strictly upper 64-blocks are read from R (ld ldr), given the inverses of
// M's diagonal 64-blocks in X's diagonal blocks (X zero below them), by
// recursive doubling on stream s.  tmp: n/2 x n doubles.
int upper_inverse_levels(mqr_handle_t hd, hipStream_t s, int64_t n, const double *R, int64_t ldr,
                         double *X, double *tmp);
// out[0:count] = sum over s < splits of P[s*stride + 0:count]; out = P when null.
// With skip, the launch does nothing when *skip is nonzero on the device.
int reduce_partials(int64_t count, double *P, int splits, int64_t stride, hipStream_t s,
                    double *out = nullptr, const int *skip = nullptr);
// Largest panel width the LDS-resident small kernels accept.
constexpr int kMaxNb = 64;
// Largest block size mqr_set_block_size accepts: above kMaxNb it is the
// outer block of two-level blocking (panels of kMaxNb, T blocks of nb).
constexpr int kMaxOuterNb = 256;
// Outer block of the default 64-column blocking (internal, T stays 64-wide).
constexpr int kDefaultOuter = 256;

}  // namespace mqr
