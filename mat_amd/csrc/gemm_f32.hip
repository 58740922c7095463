// FP32 MFMA GEMM tiles for gfx950 (v_mfma_f32_16x16x4_f32, FP32
// accumulation): the low-precision trailing update of the mixed-precision
// factorisation and the FP32 counterpart of the reference's packed
// micro-kernel behind operator* for float (Mul/__micro_kernel_packed.h:1-307).
//
// Same structure as the FP64 tiles: a TILE x TILE output tile per workgroup
// of four waves, 16-deep K slices staged in LDS with register prefetch and
// double buffering, branch-free clamped loads, split-K partial outputs.
//
// Split-BF16 variants (MODE 1 and 3) keep the FP32 operands in LDS and round
// them to BF16 when the MFMA operands are formed: x = hi + lo with
// hi = bf16(x), lo = bf16(x - hi); MODE 1 multiplies hi * hi, MODE 3 adds
// hi * lo + lo * hi (the lo * lo term, ~2^-16 relative, is dropped), all on
// gfx950's v_mfma_f32_16x16x32_bf16 (32-deep K slices in LDS) with FP32
// accumulation.
#include <algorithm>
#include <climits>

#include "mqr_internal.h"

namespace mqr {
namespace {

typedef float v4f __attribute__((ext_vector_type(4)));
typedef short v8s __attribute__((ext_vector_type(8)));
typedef __bf16 v8bf __attribute__((ext_vector_type(8)));
constexpr int NT = 256, BK = 16;

// float -> bf16 bits, round to nearest even (finite inputs)
__device__ __forceinline__ unsigned bf16_bits(float x) {
    unsigned u = __float_as_uint(x);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return u >> 16;
}

// hi/lo BF16 operands of eight FP32 values
__device__ __forceinline__ void split8(const float (&x)[8], v8bf &hi, v8bf &lo) {
    v8s h8, l8;
#pragma unroll
    for (int q = 0; q < 8; ++q) {
        const unsigned h = bf16_bits(x[q]);
        h8[q] = (short)h;
        l8[q] = (short)bf16_bits(x[q] - __uint_as_float(h << 16));
    }
    hi = __builtin_bit_cast(v8bf, h8);
    lo = __builtin_bit_cast(v8bf, l8);
}

// N consecutive floats of one column to / from memory: one vector access
// when the address is aligned to the vector, else scalar ones.  STREAM marks
// the accesses non-temporal: the output matrix C is read and written once
// per launch, and keeping it out of L2 leaves L2 to the operand Y, which
// every column of tiles reads again.
template <int N, bool STREAM = false>
__device__ __forceinline__ void load_run(float (&v)[N], const float *src) {
    typedef float vt __attribute__((ext_vector_type(N)));
    if ((reinterpret_cast<uintptr_t>(src) & (N * sizeof(float) - 1)) == 0) {
        const vt x = STREAM ? __builtin_nontemporal_load(reinterpret_cast<const vt *>(src))
                            : *reinterpret_cast<const vt *>(src);
#pragma unroll
        for (int i = 0; i < N; ++i) v[i] = x[i];
    } else {
#pragma unroll
        for (int i = 0; i < N; ++i) v[i] = STREAM ? __builtin_nontemporal_load(src + i) : src[i];
    }
}
template <int N, bool STREAM = false>
__device__ __forceinline__ void store_run(float *dst, const float (&v)[N]) {
    typedef float vt __attribute__((ext_vector_type(N)));
    if ((reinterpret_cast<uintptr_t>(dst) & (N * sizeof(float) - 1)) == 0) {
        vt x;
#pragma unroll
        for (int i = 0; i < N; ++i) x[i] = v[i];
        if (STREAM)
            __builtin_nontemporal_store(x, reinterpret_cast<vt *>(dst));
        else
            *reinterpret_cast<vt *>(dst) = x;
    } else {
#pragma unroll
        for (int i = 0; i < N; ++i) {
            if (STREAM)
                __builtin_nontemporal_store(v[i], dst + i);
            else
                dst[i] = v[i];
        }
    }
}

template <int TILE, bool KMAJOR, int BKT = BK>
__device__ __forceinline__ void slot(int tid, int q, int &r, int &k) {
    constexpr int NV = TILE * BKT / NT;
    if (!KMAJOR) {
        constexpr int RT = TILE / 2;
        r = (tid % RT) * 2 + (q & 1);
        k = (tid / RT) + (NT / RT) * (q >> 1);
    } else {
        constexpr int KT = BKT / NV;
        r = tid / KT;
        k = (tid % KT) * NV + q;
    }
}

template <int TILE, bool KMAJOR, int BKT = BK>
__device__ __forceinline__ void fetch(const float *__restrict__ X, int64_t ld, int r0, int k0,
                                      int R, int kend, int tid, float *v) {
    constexpr int NV = TILE * BKT / NT;
    if (KMAJOR) {
        // a thread's NV values are consecutive k of one row: one vector load
        // when the row and the whole run are in range
        int r, k;
        slot<TILE, true, BKT>(tid, 0, r, k);
        r += r0;
        k += k0;
        if (r < R && k + NV <= kend) {
            float t[NV];
            load_run<NV>(t, X + (int64_t)k + (int64_t)r * ld);
#pragma unroll
            for (int q = 0; q < NV; ++q) v[q] = t[q];
            return;
        }
    } else {
        // values q and q + 1 are rows r and r + 1 at the same k: one
        // two-element load when both rows and k are in range
        bool fast = true;
#pragma unroll
        for (int q = 0; q < NV; q += 2) {
            int r, k;
            slot<TILE, false, BKT>(tid, q, r, k);
            fast = fast && (r0 + r + 1 < R) && (k0 + k < kend);
        }
        if (fast) {
#pragma unroll
            for (int q = 0; q < NV; q += 2) {
                int r, k;
                slot<TILE, false, BKT>(tid, q, r, k);
                float t[2];
                load_run<2>(t, X + (int64_t)(r0 + r) + (int64_t)(k0 + k) * ld);
                v[q] = t[0];
                v[q + 1] = t[1];
            }
            return;
        }
    }
#pragma unroll
    for (int q = 0; q < NV; ++q) {
        int r, k;
        slot<TILE, KMAJOR, BKT>(tid, q, r, k);
        r += r0;
        k += k0;
        const int rr = r < R ? r : R - 1, kc = k < kend ? k : kend - 1;
        v[q] = KMAJOR ? X[(int64_t)kc + (int64_t)rr * ld] : X[(int64_t)rr + (int64_t)kc * ld];
    }
}

template <int TILE, bool KMAJOR, int BKT = BK>
__device__ __forceinline__ void stash(float (*S)[TILE + 4], int tid, const float *v, int r0, int k0,
                                      int R, int kend, float scale) {
    constexpr int NV = TILE * BKT / NT;
#pragma unroll
    for (int q = 0; q < NV; ++q) {
        int r, k;
        slot<TILE, KMAJOR, BKT>(tid, q, r, k);
        const bool in = (r0 + r < R) && (k0 + k < kend);
        S[k][r] = scale * __int_as_float(__float_as_int(v[q]) & (in ? ~0 : 0));
    }
}

template <int TILE, bool TA, bool TB, int MODE>
__global__ __launch_bounds__(NT, 2) void k_gemm_f32(GemmF32 gin, int kchunk) {
    // K slice depth: 16 for the FP32 MFMA (four K = 4 steps), 32 for the
    // BF16 modes (one v_mfma_f32_16x16x32_bf16 per tile and term)
    constexpr int BKM = MODE == 0 ? BK : 2 * BK;
    constexpr int NV = TILE * BKM / NT, BW = TILE / 32, HALF = TILE / 2;
    __shared__ float As[2][BKM][TILE + 4];
    __shared__ float Bs[2][BKM][TILE + 4];
    // tile z = batch index * splits + split index
    GemmF32 g = gin;
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
    const int kb = split * kchunk;
    const int ke = min(g.K, kb + kchunk);

    // accumulator element (i, j, r): row m0 + wm HALF + BW (l & 15) + i,
    // column n0 + wn HALF + 16 j + 4 (l >> 4) + r (swapped operands; see the
    // epilogue).  The A operand's rows are assigned so that a lane's BW row
    // tiles are BW consecutive rows: C moves as one vector per lane and
    // column.  With beta != 0 it starts as beta C, loaded before the K loop
    // so that the loads overlap the first slices; alpha is applied to op(A)
    // in LDS.
    const int arow = wm * HALF + BW * (lane & 15);  // first of the lane's rows
    const bool interior = (m0 + TILE <= g.M) && (n0 + TILE <= g.N);
    v4f acc[BW][BW];
    if (g.splits == 1 && g.beta != 0.f) {
#pragma unroll
        for (int j = 0; j < BW; ++j)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int gn = n0 + wn * HALF + j * 16 + 4 * (lane >> 4) + r;
                const int cn = gn < g.N ? gn : g.N - 1;
                const float *Cb = g.Cin ? g.Cin : g.C;
                const int64_t ldb_c = g.Cin ? g.ldcin : g.ldc;
                const float *src = Cb + (int64_t)(m0 + arow) + (int64_t)cn * ldb_c;
                float v[BW];
                if (interior) {
                    load_run<BW, true>(v, src);
                } else {
#pragma unroll
                    for (int i = 0; i < BW; ++i) {
                        const int gm = m0 + arow + i;
                        v[i] = Cb[(int64_t)(gm < g.M ? gm : g.M - 1) + (int64_t)cn * ldb_c];
                    }
                }
#pragma unroll
                for (int i = 0; i < BW; ++i) acc[i][j][r] = g.beta * v[i];
            }
    } else {
#pragma unroll
        for (int i = 0; i < BW; ++i)
#pragma unroll
            for (int j = 0; j < BW; ++j) acc[i][j] = (v4f){0.f, 0.f, 0.f, 0.f};
    }

    // one slice of the FP32 mode: four K = 4 steps of BW x BW MFMAs
    auto mma_f32 = [&](int bf) {
#pragma unroll
        for (int kk = 0; kk < BK; kk += 4) {
            float a[BW], b[BW];
#pragma unroll
            for (int i = 0; i < BW; ++i) a[i] = As[bf][kk + (lane >> 4)][arow + i];
#pragma unroll
            for (int j = 0; j < BW; ++j)
                b[j] = Bs[bf][kk + (lane >> 4)][wn * HALF + j * 16 + (lane & 15)];
#pragma unroll
            for (int i = 0; i < BW; ++i)
#pragma unroll
                for (int j = 0; j < BW; ++j)
                    acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(b[j], a[i], acc[i][j], 0, 0, 0);
        }
    };

    if constexpr (TILE == 64 && MODE == 0) {
        // The narrow FP32 tile spends 16 MFMAs per wave on a slice, less
        // than the latency of the loads that bring the next one: two slices
        // are kept in flight, in two register sets that trade roles every
        // slice.  At the top of the loop slice k0 is in As/Bs[buf], slice
        // k0 + BK (when there is one) in wa/wb, and va/vb are free.
        if (kb < ke) {
            float va[NV], vb[NV], wa[NV], wb[NV];
            fetch<TILE, TA, BKM>(g.A, g.lda, m0, kb, g.M, ke, tid, va);
            fetch<TILE, !TB, BKM>(g.B, g.ldb, n0, kb, g.N, ke, tid, vb);
            if (kb + BK < ke) {
                fetch<TILE, TA, BKM>(g.A, g.lda, m0, kb + BK, g.M, ke, tid, wa);
                fetch<TILE, !TB, BKM>(g.B, g.ldb, n0, kb + BK, g.N, ke, tid, wb);
            }
            stash<TILE, TA, BKM>(As[0], tid, va, m0, kb, g.M, ke, g.alpha);
            stash<TILE, !TB, BKM>(Bs[0], tid, vb, n0, kb, g.N, ke, 1.f);
            __syncthreads();
            int buf = 0;
            for (int k0 = kb; k0 < ke; k0 += 2 * BK) {
                if (k0 + 2 * BK < ke) {
                    fetch<TILE, TA, BKM>(g.A, g.lda, m0, k0 + 2 * BK, g.M, ke, tid, va);
                    fetch<TILE, !TB, BKM>(g.B, g.ldb, n0, k0 + 2 * BK, g.N, ke, tid, vb);
                }
                mma_f32(buf);
                if (k0 + BK >= ke) break;
                stash<TILE, TA, BKM>(As[buf ^ 1], tid, wa, m0, k0 + BK, g.M, ke, g.alpha);
                stash<TILE, !TB, BKM>(Bs[buf ^ 1], tid, wb, n0, k0 + BK, g.N, ke, 1.f);
                __syncthreads();
                buf ^= 1;
                if (k0 + 3 * BK < ke) {
                    fetch<TILE, TA, BKM>(g.A, g.lda, m0, k0 + 3 * BK, g.M, ke, tid, wa);
                    fetch<TILE, !TB, BKM>(g.B, g.ldb, n0, k0 + 3 * BK, g.N, ke, tid, wb);
                }
                mma_f32(buf);
                if (k0 + 2 * BK >= ke) break;
                stash<TILE, TA, BKM>(As[buf ^ 1], tid, va, m0, k0 + 2 * BK, g.M, ke, g.alpha);
                stash<TILE, !TB, BKM>(Bs[buf ^ 1], tid, vb, n0, k0 + 2 * BK, g.N, ke, 1.f);
                __syncthreads();
                buf ^= 1;
            }
        }
    } else if (kb < ke) {
        float va[NV], vb[NV];
        fetch<TILE, TA, BKM>(g.A, g.lda, m0, kb, g.M, ke, tid, va);
        fetch<TILE, !TB, BKM>(g.B, g.ldb, n0, kb, g.N, ke, tid, vb);
        stash<TILE, TA, BKM>(As[0], tid, va, m0, kb, g.M, ke, g.alpha);
        stash<TILE, !TB, BKM>(Bs[0], tid, vb, n0, kb, g.N, ke, 1.f);
        __syncthreads();
        int buf = 0;
        for (int k0 = kb; k0 < ke; k0 += BKM) {
            const bool more = k0 + BKM < ke;
            if (more) {
                fetch<TILE, TA, BKM>(g.A, g.lda, m0, k0 + BKM, g.M, ke, tid, va);
                fetch<TILE, !TB, BKM>(g.B, g.ldb, n0, k0 + BKM, g.N, ke, tid, vb);
            }
            if constexpr (MODE == 0) {
                mma_f32(buf);
            } else {
                // one 32-deep MFMA per tile and term: lane l supplies
                // k = 8 (l >> 4) + q, q = 0..7, of row / column l & 15 (A and
                // B pair the same k in the same lane and slot)
                v8bf ah[BW], al[BW], bh[BW], bl[BW];
#pragma unroll
                for (int i = 0; i < BW; ++i) {
                    float x[8];
#pragma unroll
                    for (int q = 0; q < 8; ++q)
                        x[q] = As[buf][8 * (lane >> 4) + q][arow + i];
                    split8(x, ah[i], al[i]);
                }
#pragma unroll
                for (int j = 0; j < BW; ++j) {
                    float x[8];
#pragma unroll
                    for (int q = 0; q < 8; ++q)
                        x[q] = Bs[buf][8 * (lane >> 4) + q][wn * HALF + j * 16 + (lane & 15)];
                    split8(x, bh[j], bl[j]);
                }
#pragma unroll
                for (int i = 0; i < BW; ++i)
#pragma unroll
                    for (int j = 0; j < BW; ++j) {
                        acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bh[j], ah[i],
                                                                            acc[i][j], 0, 0, 0);
                        if constexpr (MODE == 3) {
                            acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(
                                bh[j], al[i], acc[i][j], 0, 0, 0);
                            acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(
                                bl[j], ah[i], acc[i][j], 0, 0, 0);
                        }
                    }
            }
            if (more) {
                stash<TILE, TA, BKM>(As[buf ^ 1], tid, va, m0, k0 + BKM, g.M, ke, g.alpha);
                stash<TILE, !TB, BKM>(Bs[buf ^ 1], tid, vb, n0, k0 + BKM, g.N, ke, 1.f);
            }
            __syncthreads();
            buf ^= 1;
        }
    }
    // v_mfma_f32_16x16x4_f32 with swapped operands: lane l, register r holds
    // output row (l >> 4) * 4 + r of the transposed product, column l & 15,
    // i.e. C row = m0 + arow + i, column = base + 4 (l >> 4) + r.
    float *Cout = g.C + (g.splits > 1 ? (int64_t)split * g.split_stride : 0);
#pragma unroll
    for (int j = 0; j < BW; ++j)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int gn = n0 + wn * HALF + j * 16 + 4 * (lane >> 4) + r;
            float v[BW];
#pragma unroll
            for (int i = 0; i < BW; ++i) v[i] = acc[i][j][r];
            float *dst = Cout + (int64_t)(m0 + arow) + (int64_t)gn * g.ldc;
            if (interior) {
                store_run<BW, true>(dst, v);
            } else if (gn < g.N) {
#pragma unroll
                for (int i = 0; i < BW; ++i)
                    if (m0 + arow + i < g.M) dst[i] = v[i];
            }
        }
}

template <bool TA, bool TB, int MODE>
void launch_mode(const GemmF32 &g, hipStream_t s) {
    const int tile = (g.M >= 128 && g.N >= 128) ? 128 : 64;
    // split-K chunks start on whole slices of the mode's depth
    const int bkm = MODE == 0 ? BK : 2 * BK;
    int kchunk = (g.K + g.splits - 1) / g.splits;
    kchunk = ((kchunk + bkm - 1) / bkm) * bkm;
    dim3 grid((g.M + tile - 1) / tile, (g.N + tile - 1) / tile, g.splits * std::max(g.batch, 1));
    if (tile == 128)
        hipLaunchKernelGGL((k_gemm_f32<128, TA, TB, MODE>), grid, dim3(NT), 0, s, g, kchunk);
    else
        hipLaunchKernelGGL((k_gemm_f32<64, TA, TB, MODE>), grid, dim3(NT), 0, s, g, kchunk);
}

template <bool TA, bool TB>
void launch(const GemmF32 &g, hipStream_t s) {
    if (g.bf16_terms == 3) launch_mode<TA, TB, 3>(g, s);
    else if (g.bf16_terms == 1) launch_mode<TA, TB, 1>(g, s);
    else launch_mode<TA, TB, 0>(g, s);
}

}  // namespace

int gemm_f32(const GemmF32 &g, hipStream_t s) {
    if (g.M <= 0 || g.N <= 0) return MQR_SUCCESS;
    if (g.K <= 0) return MQR_ERR_INTERNAL;
    const double M = g.M, N = g.N, K = g.K;
    const double nb = std::max(g.batch, 1);
    ProfScope ps(g.kind, 2.0 * M * N * K * nb, 4.0 * nb * (M * K + K * N + 2.0 * M * N * g.splits),
                 s);
    if (!g.ta && !g.tb) launch<false, false>(g, s);
    else if (g.ta && !g.tb) launch<true, false>(g, s);
    else if (!g.ta && g.tb) launch<false, true>(g, s);
    else launch<true, true>(g, s);
    MQR_LAUNCH_CHECK();
    return MQR_SUCCESS;
}

}  // namespace mqr
