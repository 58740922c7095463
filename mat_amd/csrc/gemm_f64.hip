// FP64 MFMA GEMM tiles for gfx950: the device counterpart of the reference's
// packed MR x NR micro-kernel and its macro-kernel drivers
// (Mul/__micro_kernel_packed_double.h:301-397,763-817; Mul/Mul.h:239-322).
//
// Instead of packing into MR/NR panels in a separate pass
// (Mul/__pack_double.h), each workgroup stages slices of op(A) and op(B) in
// LDS, applying the unit-lower-trapezoid mask of a reflector block V on the
// fly (the reference's __pack_L_block / __pack_LT_block, Mul/__pack.h:108-189).
// Loads are branch-free (clamped in-bounds addresses, out-of-range values
// selected to zero) so that a whole slice is in flight at once.
//
// Two tiles: 64 x 64 (four waves, 2 x 2 MFMA blocks each) for small or thin
// products, and 128 x 128 with a register-prefetched double-buffered LDS
// pipeline (four waves, 4 x 4 MFMA blocks each) for the large ones.
#include <algorithm>
#include <climits>

#include "mqr_internal.h"

namespace mqr {
namespace {

typedef double v4d __attribute__((ext_vector_type(4)));

constexpr int NT = 256;

// Epilogue shared by the tiles: acc element (i, j, r) of a wave's
// (BI*16) x (BJ*16) block goes to row rowbase + i*16 + (lane & 15), column
// colbase + j*16 + (lane >> 4) + 4 r.  v_mfma_f64_16x16x4_f64 with swapped
// operands leaves output row (l >> 4) + 4 r, column l & 15 of the transposed
// product in lane l, register r: 16 consecutive rows of one column per
// row group, i.e. coalesced column-major stores.  acc already holds the
// result (alpha was applied to op(A) in LDS; beta C was loaded into acc
// before the K loop).
template <int BI, int BJ>
__device__ __forceinline__ void epilogue(const GemmF64 &g, v4d (&acc)[BI][BJ], int rowbase,
                                         int colbase, bool interior, int split) {
    const int lane = threadIdx.x & 63;
    double *Cout = g.C + (g.splits > 1 ? (int64_t)split * g.split_stride : 0);
#pragma unroll
    for (int i = 0; i < BI; ++i) {
        const int gm = rowbase + i * 16 + (lane & 15);
#pragma unroll
        for (int j = 0; j < BJ; ++j)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int gn = colbase + j * 16 + (lane >> 4) + 4 * r;
                if (interior || (gm < g.M && gn < g.N))
                    __builtin_nontemporal_store(acc[i][j][r], Cout + (int64_t)gm + (int64_t)gn * g.ldc);
            }
    }
}

// acc = beta C for the tile (clamped, always in-bounds loads), issued before
// the K loop so that the loads of C overlap the first slices.  C is read and
// written non-temporally (here and in the epilogue): each launch touches it
// once, and L2 stays with the operand every column of tiles reads again.
template <int BI, int BJ>
__device__ __forceinline__ void preload_c(const GemmF64 &g, v4d (&acc)[BI][BJ], int rowbase,
                                          int colbase) {
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int i = 0; i < BI; ++i) {
        const int gm = rowbase + i * 16 + (lane & 15);
        const int cm = gm < g.M ? gm : g.M - 1;
#pragma unroll
        for (int j = 0; j < BJ; ++j)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int gn = colbase + j * 16 + (lane >> 4) + 4 * r;
                const int cn = gn < g.N ? gn : g.N - 1;
                acc[i][j][r] =
                    g.beta * __builtin_nontemporal_load(g.Cin ? g.Cin + (int64_t)cm + (int64_t)cn * g.ldcin
                                                              : g.C + (int64_t)cm + (int64_t)cn * g.ldc);
            }
    }
}

// ---------------------------------------------------------------------------
// TILE x TILE output tile (TILE = 64 or 128), 16-deep K slices, 256 threads.
// The next slices of op(A) and op(B) are loaded into registers while the
// MFMAs consume the current LDS buffers (double buffering, one barrier per
// slice).  Each wave owns a (TILE/2) x (TILE/2) quarter as (TILE/32)^2
// 16 x 16 MFMA blocks.
// ---------------------------------------------------------------------------
constexpr int BK = 16;

// Position (r, k) of element q of a thread's share of a TILE x 16 slice of
// op(X).  KMAJOR: the stored matrix is k x r (consecutive k contiguous);
// otherwise r x k (consecutive r contiguous).
template <int TILE, bool KMAJOR>
__device__ __forceinline__ void slot(int tid, int q, int &r, int &k) {
    constexpr int NV = TILE * BK / NT;  // values per thread: 4 or 8
    if (!KMAJOR) {
        constexpr int RT = TILE / 2;     // threads along r (two values each)
        r = (tid % RT) * 2 + (q & 1);
        k = (tid / RT) + (NT / RT) * (q >> 1);
    } else {
        constexpr int KT = BK / NV;      // threads along k
        r = tid / KT;
        k = (tid % KT) * NV + q;
    }
}

// Raw loads from clamped, always in-bounds addresses.
template <int TILE, bool KMAJOR>
__device__ __forceinline__ void fetch(const double *__restrict__ X, int64_t ld, int r0, int k0,
                                      int R, int kend, int tid, double *v) {
    constexpr int NV = TILE * BK / NT;
    if (KMAJOR) {
        // a thread's NV values are consecutive k of one row: one vector load
        // when the row and the whole run are in range
        int r, k;
        slot<TILE, true>(tid, 0, r, k);
        r += r0;
        k += k0;
        if (r < R && k + NV <= kend) {
            __builtin_memcpy(v, X + (int64_t)k + (int64_t)r * ld, NV * sizeof(double));
            return;
        }
    } else {
        // values q and q + 1 are rows r and r + 1 at the same k: one
        // two-element load when both rows and k are in range
        bool fast = true;
#pragma unroll
        for (int q = 0; q < NV; q += 2) {
            int r, k;
            slot<TILE, false>(tid, q, r, k);
            fast = fast && (r0 + r + 1 < R) && (k0 + k < kend);
        }
        if (fast) {
#pragma unroll
            for (int q = 0; q < NV; q += 2) {
                int r, k;
                slot<TILE, false>(tid, q, r, k);
                __builtin_memcpy(v + q, X + (int64_t)(r0 + r) + (int64_t)(k0 + k) * ld,
                                 2 * sizeof(double));
            }
            return;
        }
    }
#pragma unroll
    for (int q = 0; q < NV; ++q) {
        int r, k;
        slot<TILE, KMAJOR>(tid, q, r, k);
        r += r0;
        k += k0;
        const int rr = r < R ? r : R - 1, kc = k < kend ? k : kend - 1;
        v[q] = KMAJOR ? X[(int64_t)kc + (int64_t)rr * ld] : X[(int64_t)rr + (int64_t)kc * ld];
    }
}

// Bounds and the reflector mask applied at LDS-store time, after the MFMAs
// of the current slice have been issued.  Stored coordinates (row, col) are
// (k, r) when KMAJOR, else (r, k).
template <int TILE, bool KMAJOR, int MASK>
__device__ __forceinline__ void stash(double (*S)[TILE + 2], int tid, const double *v, int r0,
                                      int k0, int R, int kend, int mask_cols, double scale) {
    constexpr int NV = TILE * BK / NT;
#pragma unroll
    for (int q = 0; q < NV; ++q) {
        int r, k;
        slot<TILE, KMAJOR>(tid, q, r, k);
        const int gr = r0 + r, gk = k0 + k;
        const bool in = gr < R && gk < kend;
        const int srow = KMAJOR ? gk : gr, scol = KMAJOR ? gr : gk;
        bool keep = in, one = false;
        if (MASK == MASK_VLOWER) {
            const bool m = (scol < mask_cols) && (srow <= scol);
            keep = in && !m;
            one = in && m && (srow == scol);
        }
        const unsigned long long bits = (__double_as_longlong(v[q]) & (keep ? ~0ull : 0ull)) |
                                        (one ? 0x3FF0000000000000ull : 0ull);
        S[k][r] = scale * __longlong_as_double(bits);
    }
}

template <int TILE, bool TA, bool TB, int AM, int BMK>
__global__ __launch_bounds__(NT, 2) void k_gemm_f64(GemmF64 gin, int kchunk) {
    constexpr int NV = TILE * BK / NT, BW = TILE / 32, HALF = TILE / 2;
    __shared__ double As[2][BK][TILE + 2];
    __shared__ double Bs[2][BK][TILE + 2];
    if (gin.skip_if && *gin.skip_if) return;
    // tile z = batch index * splits + split index
    GemmF64 g = gin;
    // W = V^T C style products (transposed A, few row tiles) take the
    // XCD-aware order, so that the row tiles of one column slice of C share
    // an L2; the others keep the dispatch order, in which each XCD holds the
    // same rows of op(A) for every column of tiles.
    int tx = blockIdx.x, ty = blockIdx.y, tz = blockIdx.z;
    if constexpr (TA) xcd_tile(tx, ty, tz);
    int split = tz;
    if (g.batch > 1) {
        const int bz = tz / g.splits;
        split = tz - bz * g.splits;
        g.A += bz * g.strideA;
        g.B += bz * g.strideB;
        g.C += bz * g.strideC;
        if (g.Cin) g.Cin += bz * g.strideCin;
    }
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = wave & 1, wn = wave >> 1;
    const int m0 = tx * TILE, n0 = ty * TILE;
    int kb = split * kchunk;
    int ke = min(g.K, kb + kchunk);
    // triangular operands: rows >= m0 of an upper op(A) are zero for k < m0,
    // columns < n0 + TILE of an upper op(B) are zero for k >= n0 + TILE
    // (m0 is a multiple of BK, so the slices stay aligned)
    if (g.a_upper) kb = max(kb, m0);
    if (g.b_upper) ke = min(ke, n0 + TILE);

    v4d acc[BW][BW];
    if (g.splits == 1 && g.beta != 0.0) {
        preload_c<BW, BW>(g, acc, m0 + wm * HALF, n0 + wn * HALF);
    } else {
#pragma unroll
        for (int i = 0; i < BW; ++i)
#pragma unroll
            for (int j = 0; j < BW; ++j) acc[i][j] = (v4d){0.0, 0.0, 0.0, 0.0};
    }

    if (kb < ke) {
        double va[NV], vb[NV];
        // op(A) is stored k-major when transposed; op(B)^T likewise when not
        fetch<TILE, TA>(g.A, g.lda, m0, kb, g.M, ke, tid, va);
        fetch<TILE, !TB>(g.B, g.ldb, n0, kb, g.N, ke, tid, vb);
        stash<TILE, TA, AM>(As[0], tid, va, m0, kb, g.M, ke, INT_MAX, g.alpha);
        stash<TILE, !TB, BMK>(Bs[0], tid, vb, n0, kb, g.N, ke, g.bmask_cols, 1.0);
        __syncthreads();
        int buf = 0;
        for (int k0 = kb; k0 < ke; k0 += BK) {
            const bool more = k0 + BK < ke;
            if (more) {
                fetch<TILE, TA>(g.A, g.lda, m0, k0 + BK, g.M, ke, tid, va);
                fetch<TILE, !TB>(g.B, g.ldb, n0, k0 + BK, g.N, ke, tid, vb);
            }
#pragma unroll
            for (int kk = 0; kk < BK; kk += 4) {
                double a[BW], b[BW];
#pragma unroll
                for (int i = 0; i < BW; ++i)
                    a[i] = As[buf][kk + (lane >> 4)][wm * HALF + i * 16 + (lane & 15)];
#pragma unroll
                for (int j = 0; j < BW; ++j)
                    b[j] = Bs[buf][kk + (lane >> 4)][wn * HALF + j * 16 + (lane & 15)];
#pragma unroll
                for (int i = 0; i < BW; ++i)
#pragma unroll
                    for (int j = 0; j < BW; ++j)
                        acc[i][j] =
                            __builtin_amdgcn_mfma_f64_16x16x4f64(b[j], a[i], acc[i][j], 0, 0, 0);
            }
            if (more) {
                stash<TILE, TA, AM>(As[buf ^ 1], tid, va, m0, k0 + BK, g.M, ke, INT_MAX, g.alpha);
                stash<TILE, !TB, BMK>(Bs[buf ^ 1], tid, vb, n0, k0 + BK, g.N, ke, g.bmask_cols,
                                      1.0);
            }
            __syncthreads();
            buf ^= 1;
        }
    }
    const bool interior = (m0 + TILE <= g.M) && (n0 + TILE <= g.N);
    epilogue<BW, BW>(g, acc, m0 + wm * HALF, n0 + wn * HALF, interior, split);
}

constexpr int BIG = 128, SMALL = 64;

template <bool TA, bool TB, int AM, int BMK>
void launch(const GemmF64 &g, hipStream_t s) {
    const int tile = (g.M >= BIG && g.N >= BIG && !g.small_tiles) ? BIG : SMALL;
    int kchunk = (g.K + g.splits - 1) / g.splits;
    kchunk = ((kchunk + BK - 1) / BK) * BK;
    dim3 grid((g.M + tile - 1) / tile, (g.N + tile - 1) / tile, g.splits * std::max(g.batch, 1));
    if (tile == BIG)
        hipLaunchKernelGGL((k_gemm_f64<BIG, TA, TB, AM, BMK>), grid, dim3(NT), 0, s, g, kchunk);
    else
        hipLaunchKernelGGL((k_gemm_f64<SMALL, TA, TB, AM, BMK>), grid, dim3(NT), 0, s, g, kchunk);
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
    const int t = (M >= BIG && N >= BIG) ? BIG : SMALL;
    const int tiles = ((M + t - 1) / t) * ((N + t - 1) / t);
    if (tiles >= target_blocks || K <= 512) return 1;
    int s = (target_blocks + tiles - 1) / tiles;
    const int maxs = (K + 255) / 256;
    if (s > maxs) s = maxs;
    if (s > 64) s = 64;
    return s < 1 ? 1 : s;
}

int w_pick_splits(int b, int64_t nc, int64_t h, int cus, int min_rows) {
    const int t = (b >= BIG && nc >= BIG) ? BIG : SMALL;
    const int64_t tiles = ((b + t - 1) / t) * ((nc + t - 1) / t);
    int64_t sp = (4 * (int64_t)cus + tiles - 1) / std::max<int64_t>(tiles, 1);
    sp = std::min<int64_t>(sp, (h + min_rows - 1) / min_rows);
    sp = std::min<int64_t>(sp, 256);
    return (int)std::max<int64_t>(sp, 1);
}

int gemm_f64(const GemmF64 &g, hipStream_t s) {
    if (g.M <= 0 || g.N <= 0) return MQR_SUCCESS;
    if (g.K <= 0) return MQR_ERR_INTERNAL;  // callers never ask for an empty sum
    const int key = (g.ta ? 8 : 0) | (g.tb ? 4 : 0) | (g.amask ? 2 : 0) | (g.bmask ? 1 : 0);
    const double M = g.M, N = g.N, K = g.K;
    const double out_rw = (g.splits > 1 || g.beta == 0.0) ? 1.0 : 2.0;
    const double nb = std::max(g.batch, 1);
    ProfScope ps(g.kind, 2.0 * M * N * K * nb,
                 8.0 * nb * (M * K + K * N + out_rw * M * N * g.splits), s);
    switch (key) {
        case 0: launch<false, false, MASK_NONE, MASK_NONE>(g, s); break;
        case 2: launch<false, false, MASK_VLOWER, MASK_NONE>(g, s); break;
        case 8: launch<true, false, MASK_NONE, MASK_NONE>(g, s); break;
        case 10: launch<true, false, MASK_VLOWER, MASK_NONE>(g, s); break;
        case 11: launch<true, false, MASK_VLOWER, MASK_VLOWER>(g, s); break;
        case 4: launch<false, true, MASK_NONE, MASK_NONE>(g, s); break;
        case 12: launch<true, true, MASK_NONE, MASK_NONE>(g, s); break;
        case 6: launch<false, true, MASK_VLOWER, MASK_NONE>(g, s); break;
        default: return MQR_ERR_UNSUPPORTED;
    }
    MQR_LAUNCH_CHECK();
    return MQR_SUCCESS;
}

}  // namespace mqr
