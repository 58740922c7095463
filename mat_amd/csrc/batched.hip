// Batched uneven QR: thousands of independent small matrices (m <= 512)
// factored in one persistent launch.  Per matrix the result is what
// mqr_dgeqrf gives (R on and above the diagonal, reflectors below, tau): the
// reference's blocked Householder conventions (QR/__Householder_Blocked.h:62-245),
// computed here in the precision of the data (FP32 or FP64).
//
// Scheduling: the host orders the matrices by decreasing flop count; every
// workgroup (256 threads) takes the next matrix from an atomic counter until
// the list is exhausted (longest-processing-time-first, so the big ones do
// not straggle).  The last workgroup to leave resets the counters, so the
// launch can be replayed from a hipGraph without a memset node.
//
// Per matrix, blocks of NB columns (NB = 8 or 16, per size bin from a tuned
// table):
//   panel   : the h x b panel is held in registers (thread t owns rows t and
//             t + 256) and factored column by column
//             (QR/__Householder_Blocked.h:62-84).  One workgroup reduction per
//             column gives every dot product the column needs at once: the
//             squared norm below the diagonal and x^T a_c for the columns
//             right of it (16 sums per lane, reduced across the wave by a
//             transposed butterfly, then across the four waves in LDS, with
//             one barrier).  The factored panel goes to LDS (column-major,
//             ld = h + 1) for T and the update, and back to global memory;
//             T = M^{-1}, M = strict_upper(V^T V) + diag(1 / tau) (:86-103),
//             by a recurrence run 16 lanes per column with DPP row sums;
//   update  : the trailing columns, in chunks of 16 taken by all four waves
//             together, each wave holding a quarter of the chunk's rows in
//             registers: partial W = V^T C (MFMA), the four partials summed
//             in LDS, W2 = T^T W and C -= V W2 (MFMA, from registers), so
//             the chunk is read once and written once (:143-212).
#include <algorithm>
#include <cstring>
#include <mutex>
#include <numeric>
#include <utility>
#include <vector>

#include "mqr_internal.h"

namespace mqr {

// Per-matrix descriptor, in execution order.
struct BatchDesc {
    void *A;
    void *tau;
    int64_t lda;
    int m, n;
    int nb;   // panel width: 8 or 16
    int pad;
};

namespace {

constexpr int BNT = 256;       // threads per workgroup
constexpr int BMAXM = 512;     // largest m accepted
constexpr int BLD = BMAXM + 1; // LDS leading dimension bound

typedef double v4d __attribute__((ext_vector_type(4)));
template <typename T>
using GPtr = __attribute__((address_space(1))) T *;
typedef float v4f __attribute__((ext_vector_type(4)));

template <typename T>
struct Mf;
template <>
struct Mf<double> {
    typedef v4d acc;
    static __device__ __forceinline__ acc run(double a, double b, acc c) {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
    // output row of lane l, register r (the column is l & 15)
    static __device__ __forceinline__ int row(int l, int r) { return (l >> 4) + 4 * r; }
    // physical row (within a 16-row tile) given to tile row i, chosen so that
    // lane l, register r lands on row 4 (l >> 4) + r: four consecutive rows
    static __device__ __forceinline__ int phys(int i) { return ((i & 3) << 2) | (i >> 2); }
};
template <>
struct Mf<float> {
    typedef v4f acc;
    static __device__ __forceinline__ acc run(float a, float b, acc c) {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
    static __device__ __forceinline__ int row(int l, int r) { return 4 * (l >> 4) + r; }
    static __device__ __forceinline__ int phys(int i) { return i; }
};

// Cross-lane moves of 32- and 64-bit values on the VALU (no LDS round trip).
template <typename T>
__device__ __forceinline__ uint64_t bits64(T x) {
    return sizeof(T) == 8 ? __builtin_bit_cast(uint64_t, (double)x)
                          : (uint64_t)__builtin_bit_cast(uint32_t, (float)x);
}
template <typename T>
__device__ __forceinline__ T from_bits(uint32_t lo, uint32_t hi) {
    if constexpr (sizeof(T) == 8)
        return __builtin_bit_cast(double, ((uint64_t)hi << 32) | lo);
    else
        return __builtin_bit_cast(float, lo);
}
// DPP lane move within rows of 16 lanes (quad_perm or row rotate controls)
template <int CTRL, typename T>
__device__ __forceinline__ T dpp(T x) {
    const uint64_t u = bits64(x);
    const uint32_t lo = (uint32_t)__builtin_amdgcn_mov_dpp((int)(uint32_t)u, CTRL, 0xF, 0xF, false);
    uint32_t hi = 0;
    if constexpr (sizeof(T) == 8)
        hi = (uint32_t)__builtin_amdgcn_mov_dpp((int)(uint32_t)(u >> 32), CTRL, 0xF, 0xF, false);
    return from_bits<T>(lo, hi);
}
constexpr int kQuadXor1 = 0xB1;  // quad_perm [1, 0, 3, 2]
constexpr int kQuadXor2 = 0x4E;  // quad_perm [2, 3, 0, 1]
constexpr int kRowRor4 = 0x124;  // row_ror:4
constexpr int kRowRor8 = 0x128;  // row_ror:8, the same as lane ^ 8 in a row

// a + b after v_permlane32_swap (MASK 32: lanes 32-63 of a trade places with
// lanes 0-31 of b) or v_permlane16_swap (MASK 16: odd rows of a with even
// rows of b): the low half of the lanes gets a summed over the lane pair
// (l, l ^ MASK), the high half gets b summed likewise.
template <int MASK, typename T>
__device__ __forceinline__ T swap_sum(T a, T b) {
    const uint64_t ua = bits64(a), ub = bits64(b);
    uint32_t x[2], y[2];
#pragma unroll
    for (int w = 0; w < (sizeof(T) == 8 ? 2 : 1); ++w) {
        const uint32_t pa = (uint32_t)(ua >> (32 * w)), pb = (uint32_t)(ub >> (32 * w));
        if constexpr (MASK == 32) {
            const auto r = __builtin_amdgcn_permlane32_swap(pa, pb, false, false);
            x[w] = r[0];
            y[w] = r[1];
        } else {
            const auto r = __builtin_amdgcn_permlane16_swap(pa, pb, false, false);
            x[w] = r[0];
            y[w] = r[1];
        }
    }
    if constexpr (sizeof(T) == 8)
        return from_bits<T>(x[0], x[1]) + from_bits<T>(y[0], y[1]);
    else
        return from_bits<T>(x[0], 0) + from_bits<T>(y[0], 0);
}

// Sixteen wave sums at once (a transposed butterfly): the exchanges over lane
// bits 5, 4, 3 and 2 each halve the values a lane holds, keeping the half
// that the bit selects plus the partner's copy of it; the last two steps sum
// over the four lanes left holding the same value.  Lane l returns the wave
// sum of v[l >> 2].
template <typename T>
__device__ __forceinline__ T wsum16(T (&v)[16], int lane) {
#pragma unroll
    for (int k = 0; k < 8; ++k) v[k] = swap_sum<32>(v[k], v[8 + k]);
#pragma unroll
    for (int k = 0; k < 4; ++k) v[k] = swap_sum<16>(v[k], v[4 + k]);
    {
        const bool hi = (lane & 8) != 0;
#pragma unroll
        for (int k = 0; k < 2; ++k) {
            const T keep = hi ? v[2 + k] : v[k];
            const T send = hi ? v[k] : v[2 + k];
            v[k] = keep + dpp<kRowRor8>(send);
        }
    }
    T x;
    {
        const bool hi = (lane & 4) != 0;
        const T keep = hi ? v[1] : v[0];
        const T send = hi ? v[0] : v[1];
        x = keep + __shfl_xor(send, 4, 64);
    }
    x += dpp<kQuadXor2>(x);
    x += dpp<kQuadXor1>(x);
    return x;
}

// Sum over each row of 16 lanes, returned to every lane of the row.
template <typename T>
__device__ __forceinline__ T rowsum16(T x) {
    x += dpp<kQuadXor1>(x);
    x += dpp<kQuadXor2>(x);
    x += dpp<kRowRor4>(x);
    x += dpp<kRowRor8>(x);
    return x;
}

template <typename T>
struct Shared {
    T P[16 * BLD];        // panel, column-major, ld = h + 1
    T W[4][16][17];       // per-wave partials: V^T V, then V^T C of each chunk
    T Tm[16][17];         // T of the panel
    T G[16][17];          // V^T V of the panel
    T red[2][4][16];      // per-column partial sums, double-buffered
    T rowr[2][16];        // row r of the panel while column r is reduced
    T tau[16];
    int idx;
};

// unit-lower-trapezoid V(r, p) of the b-column panel factored in LDS; zero
// outside rows [0, h) and columns [0, b)
template <typename T>
__device__ __forceinline__ T vget(const T *P, int ld, int h, int b, int r, int p) {
    const T x = (p < b && r < h) ? P[p * ld + r] : T(0);
    return r > p ? x : (r == p && p < b ? T(1) : T(0));
}

// Column R of the panel (QR/__Householder_Blocked.h:62-84).  Thread t owns
// rows t (a0) and t + BNT (a1) in registers; rows at or past h hold zeros and
// stay zero.  With x = a_R: sigma^2 = sum_{i>R} x_i^2 and
// d_c = sum_{i>R} x_i a_c(i) (c > R) come from one workgroup reduction; then
// ||x|| = sqrt(x_R^2 + sigma^2), v = x / (x_R + sign(x_R) ||x||) below the
// diagonal, tau = (x_R + sign(x_R) ||x||) / (sign(x_R) ||x||), and
// a_c -= tau (a_c(R) + v^T a_c(>R)) v, where v^T a_c(>R) = d_c / (...).
// A zero column (||x|| = 0) is left as it is, with tau = 0 (H = I).
template <int R, typename T>
__device__ __forceinline__ void panel_column(T (&a0)[16], T (&a1)[16], Shared<T> &sh, int b,
                                             int tid) {
    if (R >= b) return;  // uniform
    const int lane = tid & 63, wave = tid >> 6;
    T part[16];
#pragma unroll
    for (int c = 0; c < 16; ++c) part[c] = T(0);
    {
        const T x0 = tid > R ? a0[R] : T(0);
        const T x1 = a1[R];  // row t + BNT is always below row R
#pragma unroll
        for (int c = R; c < 16; ++c) part[c] += x0 * a0[c] + x1 * a1[c];
    }
    const T tot = wsum16(part, lane);
    // double-buffered slots: the next column writes the other buffer, and the
    // one after it only once every thread has passed the next barrier
    T *red = &sh.red[R & 1][0][0];
    T *rowr = sh.rowr[R & 1];
    if ((lane & 3) == 0) red[wave * 16 + (lane >> 2)] = tot;
    if (tid == R) {
#pragma unroll
        for (int c = R; c < 16; ++c) rowr[c] = a0[c];
    }
    __syncthreads();
    const T alpha = rowr[R];
    const T sig2 = (red[R] + red[16 + R]) + (red[32 + R] + red[48 + R]);
    const T nrm = sqrt(alpha * alpha + sig2);
    // a zero column (nrm = 0) keeps its entries, with tau = 0 (H = I): hr
    // and t are selected to zero, which makes every update below add zero
    // (branch-free, so the register panel is not copied at a join)
    const bool zero = nrm == T(0);  // uniform: every thread read the same sums
    const T cr = alpha < T(0) ? -nrm : nrm;
    const T w0 = alpha + cr;
    const T hr = zero ? T(0) : T(1) / w0;
    const T t = zero ? T(0) : w0 / cr;
    if (tid == 0) sh.tau[R] = t;
    // the reflector's entry for each owned row: v_i below the diagonal, 1 on
    // it, 0 above (rows above R are final and get a zero update)
    const T v0 = a0[R] * hr;
    const T vr0 = tid > R ? v0 : (tid == R ? T(1) : T(0));
    a0[R] = zero ? a0[R] : (tid > R ? v0 : (tid == R ? -cr : a0[R]));
    const T vr1 = a1[R] * hr;
    a1[R] = zero ? a1[R] : vr1;
#pragma unroll
    for (int c = R + 1; c < 16; ++c) {
        const T dc = (red[c] + red[16 + c]) + (red[32 + c] + red[48 + c]);
        const T cf = -t * (rowr[c] + hr * dc);
        a0[c] += cf * vr0;
        a1[c] += cf * vr1;
        // bound the partial sums in flight (register pressure)
        if ((c & 3) == 3) __builtin_amdgcn_sched_barrier(0);
    }
}

template <typename T, int... R>
__device__ __forceinline__ void panel_columns(T (&a0)[16], T (&a1)[16], Shared<T> &sh, int b,
                                              int tid, std::integer_sequence<int, R...>) {
    (panel_column<R>(a0, a1, sh, b, tid), ...);
}

template <typename T>
__device__ __forceinline__ void factor_one(const BatchDesc &d, Shared<T> &sh) {
    // the descriptor's pointers are global memory: global (not flat)
    // loads and stores, which do not hold up the LDS wait counter
    GPtr<T> A = (GPtr<T>)d.A;
    GPtr<T> tau = (GPtr<T>)d.tau;
    const int64_t lda = d.lda;
    const int m = d.m, n = d.n, k = min(m, n), NB = d.nb;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    for (int j = 0; j < k; j += NB) {
        const int b = min(NB, k - j);
        const int h = m - j;
        const int ld = h + 1;
        GPtr<T> Aj = A + j + (int64_t)j * lda;
        // ---- panel to registers: thread t owns rows t and t + BNT ----
        T a0[16], a1[16];
#pragma unroll
        for (int c = 0; c < 16; ++c) {
            a0[c] = (tid < h && c < b) ? Aj[tid + (int64_t)c * lda] : T(0);
            a1[c] = (tid + BNT < h && c < b) ? Aj[tid + BNT + (int64_t)c * lda] : T(0);
        }
        // ---- column-by-column reflectors (QR/__Householder_Blocked.h:62-84) ----
        panel_columns(a0, a1, sh, b, tid, std::make_integer_sequence<int, 16>());
        // ---- factored panel to LDS (for T and the update) ----
#pragma unroll
        for (int c = 0; c < 16; ++c)
            if (c < b) {
                if (tid < h) sh.P[c * ld + tid] = a0[c];
                if (tid + BNT < h) sh.P[c * ld + tid + BNT] = a1[c];
            }
        __syncthreads();
        // ---- T from G = V^T V: T^{-1} = strict_upper(G) + diag(1 / tau) ----
        // G on the MFMA: wave w sums rows [w hq, (w + 1) hq), then the four
        // partial Gram matrices are added
        {
            typedef typename Mf<T>::acc acc_t;
            const int hq = (((h + 3) / 4) + 3) & ~3;
            acc_t g = {0, 0, 0, 0};
            for (int r0 = wave * hq; r0 < min(h, (wave + 1) * hq); r0 += 4) {
                const T v = vget(sh.P, ld, h, b, r0 + (lane >> 4), lane & 15);
                g = Mf<T>::run(v, v, g);
            }
#pragma unroll
            for (int rr = 0; rr < 4; ++rr) sh.W[wave][Mf<T>::row(lane, rr)][lane & 15] = g[rr];
        }
        __syncthreads();
        {
            const int p = tid & 15, q = tid >> 4;
            sh.G[p][q] = (sh.W[0][p][q] + sh.W[1][p][q]) + (sh.W[2][p][q] + sh.W[3][p][q]);
        }
        __syncthreads();
        {
            // T = M^{-1} by T(i, c) = tau_i (delta_ic - sum_{i<l<=c} G(i, l) T(l, c)),
            // i from 15 down.  Wave w builds columns c = 4 w + (lane >> 4);
            // lane l = lane & 15 of a row holds T(l, c) and column l of G, and
            // the sum over l is a row sum.  tau_i = 0 (an identity reflector,
            // or i >= b) leaves row and column i of T zero, which is what
            // removing the reflector from M gives.
            const int l = lane & 15, c = 4 * wave + (lane >> 4);
            T g[16], tv[16];
#pragma unroll
            for (int i = 0; i < 16; ++i) {
                g[i] = sh.G[i][l];
                tv[i] = i < b ? sh.tau[i] : T(0);
            }
            T x = T(0);
#pragma unroll
            for (int i = 15; i >= 0; --i) {
                const T sum = rowsum16(l > i ? g[i] * x : T(0));
                const T xi = tv[i] * ((i == c ? T(1) : T(0)) - sum);
                if (l == i) x = xi;
            }
            sh.Tm[l][c] = x;
        }
        // ---- panel back to global (from LDS: the register copy is dead), tau ----
        for (int e = tid; e < b * h; e += BNT) {
            const int r = e % h, c = e / h;
            Aj[r + (int64_t)c * lda] = sh.P[c * ld + r];
        }
        if (tid < b) tau[j + tid] = sh.tau[tid];
        __syncthreads();
        // ---- trailing update, one 16-column chunk at a time by all four
        // waves: wave w holds rows [w hs, (w + 1) hs) of the chunk in
        // registers (hs <= 128) from W = V^T C until C -= V (T^T W) is stored,
        // so C is read once and written once per panel (QR/__Householder_Blocked.h:143-212).
        // Register layout of a 16-row tile (the MFMA's B operand and output):
        // lane l, register u holds row 4 (l >> 4) + u, column l & 15, the
        // MFMA's row index Mf::row(l, u) mapped through Mf::phys.  The four
        // partial W are summed in LDS (two barriers per chunk); every wave then
        // forms W2 = T^T W and V W2 on the MFMA from registers.
        typedef typename Mf<T>::acc acc_t;
        const int c0base = j + b;
        const int nchunks = (n - c0base + 15) / 16;
        const int hs = 16 * ((h + 63) / 64);  // 4 hs >= h, multiple of 16
        const int r0w = wave * hs;
        const int ntile = max(0, min(hs, h - r0w) + 15) / 16;
        T tA[4];  // T^T as the A operand of W2 = T^T W
#pragma unroll
        for (int kk = 0; kk < 4; ++kk) tA[kk] = sh.Tm[Mf<T>::row(lane, kk)][lane & 15];
        for (int q = 0; q < nchunks; ++q) {
            const int c0 = c0base + 16 * q;
            const int col = c0 + (lane & 15);
            const bool cin = col < n;
            GPtr<T> Ccol = A + j + (int64_t)(cin ? col : c0) * lda;  // row j of the column
            T cr[8][4];
            // lane l, register u: row 4 (l >> 4) + u of tile t, four
            // consecutive rows of one column per lane (one vector load)
            const int lr = 4 * (lane >> 4);
#pragma unroll
            for (int t = 0; t < 8; ++t) {
                const int row = r0w + 16 * t + lr;
                if (t < ntile && cin && row + 3 < h) {
                    __builtin_memcpy(&cr[t][0], Ccol + row, 4 * sizeof(T));
                } else {
#pragma unroll
                    for (int u = 0; u < 4; ++u)
                        cr[t][u] = (t < ntile && cin && row + u < h) ? Ccol[row + u] : T(0);
                }
            }
            // partial W = V^T C over this wave's rows, two accumulators
            acc_t wa = {0, 0, 0, 0}, wb = {0, 0, 0, 0};
#pragma unroll
            for (int t = 0; t < 8; ++t) {
                if (t < ntile) {
#pragma unroll
                    for (int u = 0; u < 4; u += 2) {
                        const int row = r0w + 16 * t + lr + u;
                        wa = Mf<T>::run(vget(sh.P, ld, h, b, row, lane & 15), cr[t][u], wa);
                        wb = Mf<T>::run(vget(sh.P, ld, h, b, row + 1, lane & 15), cr[t][u + 1], wb);
                    }
                }
            }
            __syncthreads();  // the previous chunk's partials have been read
#pragma unroll
            for (int rr = 0; rr < 4; ++rr)
                sh.W[wave][Mf<T>::row(lane, rr)][lane & 15] = wa[rr] + wb[rr];
            __syncthreads();
            acc_t w;
#pragma unroll
            for (int rr = 0; rr < 4; ++rr) {
                const int r = Mf<T>::row(lane, rr), cc = lane & 15;
                w[rr] = (sh.W[0][r][cc] + sh.W[1][r][cc]) + (sh.W[2][r][cc] + sh.W[3][r][cc]);
            }
            // W2 = T^T W (K = 16: four MFMAs, W's registers are the B operand)
            acc_t w2 = {0, 0, 0, 0};
#pragma unroll
            for (int kk = 0; kk < 4; ++kk) w2 = Mf<T>::run(tA[kk], w[kk], w2);
            // C -= V W2, tile by tile, and store
#pragma unroll
            for (int t = 0; t < 8; ++t) {
                if (t < ntile) {
                    acc_t d = {0, 0, 0, 0};
                    const int vrow = r0w + 16 * t + Mf<T>::phys(lane & 15);
#pragma unroll
                    for (int kk = 0; kk < 4; ++kk)
                        d = Mf<T>::run(vget(sh.P, ld, h, b, vrow, Mf<T>::row(lane, kk)), w2[kk], d);
                    const int row = r0w + 16 * t + lr;
#pragma unroll
                    for (int u = 0; u < 4; ++u) cr[t][u] -= d[u];
                    if (cin && row + 3 < h) {
                        __builtin_memcpy(Ccol + row, &cr[t][0], 4 * sizeof(T));
                    } else {
#pragma unroll
                        for (int u = 0; u < 4; ++u)
                            if (cin && row + u < h) Ccol[row + u] = cr[t][u];
                    }
                }
            }
        }
        __syncthreads();
    }
}

template <typename T>
__device__ __forceinline__ void qr_batched_body(const BatchDesc *__restrict__ desc, int batch,
                                                int *__restrict__ counters) {
    // static LDS: every address in the layout is a compile-time offset
    __shared__ Shared<T> sh;
    for (;;) {
        if (threadIdx.x == 0) sh.idx = atomicAdd(&counters[0], 1);
        __syncthreads();
        const int i = sh.idx;
        __syncthreads();
        if (i >= batch) break;
        factor_one<T>(desc[i], sh);
    }
    // the last workgroup out resets the counters for the next replay
    if (threadIdx.x == 0) {
        __threadfence();
        const int done = atomicAdd(&counters[1], 1);
        if (done == (int)gridDim.x - 1) {
            atomicExch(&counters[0], 0);
            atomicExch(&counters[1], 0);
        }
    }
}

// Register budgets: FP64 fits two waves per SIMD (LDS allows two workgroups
// per CU), FP32 three.
template <typename T>
__global__ void k_qr_batched(const BatchDesc *__restrict__ desc, int batch, int *__restrict__ counters);
template <>
__global__ __launch_bounds__(BNT, 2) void k_qr_batched<double>(const BatchDesc *__restrict__ desc,
                                                               int batch,
                                                               int *__restrict__ counters) {
    qr_batched_body<double>(desc, batch, counters);
}
template <>
__global__ __launch_bounds__(BNT, 3) void k_qr_batched<float>(const BatchDesc *__restrict__ desc,
                                                              int batch,
                                                              int *__restrict__ counters) {
    qr_batched_body<float>(desc, batch, counters);
}

}  // namespace

// Panel widths per m bin (<= 64, <= 128, <= 256, <= 512) from the
// autotuner (mat_amd/autotune_batched.py), kTunedNb[double][bin].
#include "tuned_batched.h"

int batched_nb(bool f64, int m, int tile) {
    if (tile == 0) return 8;
    if (tile == 1) return 16;
    const int bin = m <= 64 ? 0 : m <= 128 ? 1 : m <= 256 ? 2 : 3;
    return kTunedNb[f64 ? 1 : 0][bin];
}

// The launch's schedule: resident workgroups per CU (0: as many as the
// kernel's resources allow) and the order of the queue (0: decreasing flop
// count, 1: as given, 2: alternately the largest and the smallest left).
struct BatchedSched {
    int per_cu = -1;  // -1: the tuned value
    int order = -1;   // -1: the tuned value
};

struct BatchedPlan {
    bool f64 = true;
    int batch = 0;
    BatchDesc *d_desc = nullptr;
    int *d_counters = nullptr;
    unsigned grid = 0;
    size_t smem = 0;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
};

static int batched_setup(mqr_handle_t hd, bool f64, int64_t batch, const int32_t *m,
                         const int32_t *n, const int64_t *lda, void *const *A, void *const *tau,
                         int tile, BatchedPlan &p, BatchedSched sched = BatchedSched()) {
    const int qorder = sched.order >= 0 ? sched.order : kTunedOrder[f64 ? 1 : 0];
    int req_cu = sched.per_cu >= 0 ? sched.per_cu : kTunedPerCu[f64 ? 1 : 0];
    std::vector<int> order(batch);
    std::iota(order.begin(), order.end(), 0);
    auto cost = [&](int i) {
        const double mm = m[i], nn = std::min(m[i], n[i]);
        return 2.0 * mm * nn * n[i] - 2.0 * nn * nn * nn / 3.0 + mm;
    };
    if (qorder != 1)
        std::stable_sort(order.begin(), order.end(),
                         [&](int a, int b) { return cost(a) > cost(b); });
    if (qorder == 2) {  // largest, smallest, next largest, next smallest, ...
        std::vector<int> alt;
        alt.reserve(batch);
        for (int64_t lo = 0, hi = batch - 1; lo <= hi; ++lo, --hi) {
            alt.push_back(order[lo]);
            if (lo != hi) alt.push_back(order[hi]);
        }
        order.swap(alt);
    }
    std::vector<BatchDesc> h(batch);
    for (int64_t e = 0; e < batch; ++e) {
        const int i = order[e];
        h[e].A = A[i];
        h[e].tau = tau[i];
        h[e].lda = lda[i];
        h[e].m = m[i];
        h[e].n = n[i];
        h[e].nb = batched_nb(f64, m[i], tile);
        h[e].pad = 0;
    }
    p.f64 = f64;
    p.batch = (int)batch;
    MQR_HIP(hipMalloc(&p.d_desc, sizeof(BatchDesc) * std::max<int64_t>(batch, 1)));
    MQR_HIP(hipMalloc(&p.d_counters, 2 * sizeof(int)));
    MQR_HIP(hipMemcpyAsync(p.d_desc, h.data(), sizeof(BatchDesc) * batch, hipMemcpyHostToDevice,
                           hd->stream));
    MQR_HIP(hipMemsetAsync(p.d_counters, 0, 2 * sizeof(int), hd->stream));
    MQR_HIP(hipStreamSynchronize(hd->stream));
    // the kernel's LDS is static (sizeof(Shared<T>)); no dynamic LDS
    p.smem = 0;
    const void *fn = f64 ? (const void *)k_qr_batched<double> : (const void *)k_qr_batched<float>;
    int per_cu = 0;
    MQR_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, BNT, p.smem));
    per_cu = std::max(per_cu, 1);
    if (req_cu > 0) per_cu = std::min(per_cu, req_cu);
    p.grid = (unsigned)std::min<int64_t>(batch, (int64_t)per_cu * hd->cu_count);
    return MQR_SUCCESS;
}

static int batched_launch(const BatchedPlan &p, hipStream_t s) {
    if (p.batch == 0) return MQR_SUCCESS;
    if (p.f64)
        hipLaunchKernelGGL(k_qr_batched<double>, dim3(p.grid), dim3(BNT), p.smem, s, p.d_desc,
                           p.batch, p.d_counters);
    else
        hipLaunchKernelGGL(k_qr_batched<float>, dim3(p.grid), dim3(BNT), p.smem, s, p.d_desc,
                           p.batch, p.d_counters);
    MQR_LAUNCH_CHECK();
    return MQR_SUCCESS;
}

static void batched_free(BatchedPlan &p) {
    if (p.exec) (void)hipGraphExecDestroy(p.exec);
    if (p.graph) (void)hipGraphDestroy(p.graph);
    if (p.d_desc) (void)hipFree(p.d_desc);
    if (p.d_counters) (void)hipFree(p.d_counters);
    p = BatchedPlan();
}

int geqrf_batched(mqr_handle_t hd, bool f64, int64_t batch, const int32_t *m, const int32_t *n,
                  const int64_t *lda, void *const *A, void *const *tau, int tile) {
    BatchedPlan p;
    int st = batched_setup(hd, f64, batch, m, n, lda, A, tau, tile, p);
    if (st == MQR_SUCCESS) {
        ProfScope ps(K_PANEL_SMALL, 0.0, 0.0, hd->stream);
        st = batched_launch(p, hd->stream);
    }
    if (st == MQR_SUCCESS && hipStreamSynchronize(hd->stream) != hipSuccess) st = MQR_ERR_HIP;
    batched_free(p);
    return st;
}

}  // namespace mqr

struct mqr_batched_plan_st {
    mqr::BatchedPlan p;
    int device = 0;
};

namespace mqr {

int batched_plan_create(mqr_handle_t hd, bool f64, int64_t batch, const int32_t *m,
                        const int32_t *n, const int64_t *lda, void *const *A, void *const *tau,
                        int tile, mqr_batched_plan_st **out, int per_cu, int order) {
    mqr_batched_plan_st *pl = new (std::nothrow) mqr_batched_plan_st();
    if (!pl) return MQR_ERR_ALLOC;
    pl->device = hd->device;
    BatchedSched sched;
    sched.per_cu = per_cu;
    sched.order = order;
    int st = batched_setup(hd, f64, batch, m, n, lda, A, tau, tile, pl->p, sched);
    if (st == MQR_SUCCESS && batch > 0) {
        // capture the single persistent launch; replays need no memset node
        // because the kernel resets its own counters
        hipStream_t cs = nullptr;
        if (hipStreamCreateWithFlags(&cs, hipStreamNonBlocking) != hipSuccess) st = MQR_ERR_HIP;
        if (st == MQR_SUCCESS &&
            hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal) != hipSuccess)
            st = MQR_ERR_HIP;
        if (st == MQR_SUCCESS) {
            const int ls = batched_launch(pl->p, cs);
            const hipError_t ec = hipStreamEndCapture(cs, &pl->p.graph);
            if (ls != MQR_SUCCESS) st = ls;
            else if (ec != hipSuccess) st = MQR_ERR_HIP;
        }
        if (st == MQR_SUCCESS &&
            hipGraphInstantiate(&pl->p.exec, pl->p.graph, nullptr, nullptr, 0) != hipSuccess)
            st = MQR_ERR_HIP;
        if (cs) (void)hipStreamDestroy(cs);
        if (st != MQR_SUCCESS) (void)hipGetLastError();
    }
    if (st != MQR_SUCCESS) {
        batched_free(pl->p);
        delete pl;
        return st;
    }
    *out = pl;
    return MQR_SUCCESS;
}

int batched_plan_run(mqr_handle_t hd, mqr_batched_plan_st *pl) {
    if (pl->p.batch == 0) return MQR_SUCCESS;
    ProfScope ps(K_PANEL_SMALL, 0.0, 0.0, hd->stream);
    MQR_HIP(hipGraphLaunch(pl->p.exec, hd->stream));
    return MQR_SUCCESS;
}

void batched_plan_destroy(mqr_batched_plan_st *pl) {
    batched_free(pl->p);
    delete pl;
}

int batched_plan_device(const mqr_batched_plan_st *pl) { return pl->device; }

}  // namespace mqr

// ---- allocation-free variant ----------------------------------------------
// mqr_geqrf_batched_uneven_noalloc runs the same persistent kernel with its
// descriptor queue and counters in the fixed device array below instead of
// per-call allocations.  Every call on a GPU shares that GPU's copy of the
// array, so calls that overlap on one GPU race; a process-wide mutex
// serialises the calls of one process (each call blocks until its kernel is
// done), and each call clears the counters before its launch.
namespace mqr {
namespace {

constexpr size_t kNoAllocDescBytes = sizeof(BatchDesc) * (size_t)MQR_NOALLOC_MAX_BATCH;
constexpr size_t kNoAllocBytes = kNoAllocDescBytes + 256;  // + the two queue counters
constexpr int kMaxDevices = 64;

// One copy per GPU, alive until the process ends.  Host code never passes
// this name to a kernel or a copy: it uses the per-device address that
// hipGetSymbolAddress returned (g_scratch_ptr).
static __device__ __align__(256) unsigned char scratch[kNoAllocBytes];

struct NoAllocDevice {
    unsigned char *ptr = nullptr;  // device address of `scratch` on this GPU
    unsigned grid64 = 0, grid32 = 0;
};
NoAllocDevice g_scratch_ptr[kMaxDevices];
bool g_scratch_ready[kMaxDevices] = {};
std::mutex g_noalloc_mutex;
// host staging of the descriptors (static: no allocation at call time)
BatchDesc g_stage[MQR_NOALLOC_MAX_BATCH];

__global__ void k_clear_counters(int *c) {
    if (threadIdx.x < 2) c[threadIdx.x] = 0;
}

}  // namespace

// Looks up this device's address of the scratch array (the device must be
// current) and the persistent grids; once per device, from mqr_create.
int batched_noalloc_init(int device) {
    if (device < 0 || device >= kMaxDevices) return MQR_ERR_NO_DEVICE;
    std::lock_guard<std::mutex> lock(g_noalloc_mutex);
    if (g_scratch_ready[device]) return MQR_SUCCESS;
    void *p = nullptr;
    MQR_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) return MQR_ERR_INTERNAL;
    int cus = 0, pc64 = 0, pc32 = 0;
    MQR_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device));
    MQR_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&pc64, (const void *)k_qr_batched<double>,
                                                         BNT, 0));
    MQR_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&pc32, (const void *)k_qr_batched<float>,
                                                         BNT, 0));
    NoAllocDevice &d = g_scratch_ptr[device];
    d.ptr = (unsigned char *)p;
    // residency as tuned for the allocating path (0: the occupancy limit)
    if (kTunedPerCu[1] > 0) pc64 = std::min(pc64, kTunedPerCu[1]);
    if (kTunedPerCu[0] > 0) pc32 = std::min(pc32, kTunedPerCu[0]);
    d.grid64 = (unsigned)(std::max(pc64, 1) * std::max(cus, 1));
    d.grid32 = (unsigned)(std::max(pc32, 1) * std::max(cus, 1));
    g_scratch_ready[device] = true;
    return MQR_SUCCESS;
}

// The batched factorisation without any device or host allocation: the
// descriptors are built in static host memory, copied into this GPU's
// scratch, the counters cleared by a kernel, and the persistent kernel run;
// returns once it is done.  batch <= MQR_NOALLOC_MAX_BATCH (checked by the
// caller).
int geqrf_batched_noalloc(mqr_handle_t hd, bool f64, int64_t batch, const int32_t *m,
                          const int32_t *n, const int64_t *lda, void *const *A, void *const *tau,
                          int tile) {
    if (batch > MQR_NOALLOC_MAX_BATCH) return MQR_ERR_TOO_LARGE;
    if (hd->device < 0 || hd->device >= kMaxDevices) return MQR_ERR_NO_DEVICE;
    std::lock_guard<std::mutex> lock(g_noalloc_mutex);
    if (!g_scratch_ready[hd->device]) return MQR_ERR_NOT_INITIALIZED;
    const NoAllocDevice &d = g_scratch_ptr[hd->device];
    for (int64_t i = 0; i < batch; ++i) {
        BatchDesc &e = g_stage[i];
        e.A = A[i];
        e.tau = tau[i];
        e.lda = lda[i];
        e.m = m[i];
        e.n = n[i];
        e.nb = batched_nb(f64, m[i], tile);
        e.pad = 0;
    }
    // longest first, by the same flop count as the allocating path (the
    // order only changes which workgroup takes a matrix, not its factors)
    auto cost = [](const BatchDesc &x) {
        const double mm = x.m, nn = std::min(x.m, x.n);
        return 2.0 * mm * nn * x.n - 2.0 * nn * nn * nn / 3.0 + mm;
    };
    std::sort(g_stage, g_stage + batch,
              [&](const BatchDesc &a, const BatchDesc &b) { return cost(a) > cost(b); });
    BatchDesc *ddesc = (BatchDesc *)d.ptr;
    int *dcount = (int *)(d.ptr + kNoAllocDescBytes);
    MQR_HIP(hipMemcpyAsync(ddesc, g_stage, sizeof(BatchDesc) * (size_t)batch,
                           hipMemcpyHostToDevice, hd->stream));
    hipLaunchKernelGGL(k_clear_counters, dim3(1), dim3(64), 0, hd->stream, dcount);
    MQR_LAUNCH_CHECK();
    const unsigned grid = (unsigned)std::min<int64_t>(batch, f64 ? d.grid64 : d.grid32);
    if (f64)
        hipLaunchKernelGGL(k_qr_batched<double>, dim3(grid), dim3(BNT), 0, hd->stream,
                           (const BatchDesc *)ddesc, (int)batch, dcount);
    else
        hipLaunchKernelGGL(k_qr_batched<float>, dim3(grid), dim3(BNT), 0, hd->stream,
                           (const BatchDesc *)ddesc, (int)batch, dcount);
    MQR_LAUNCH_CHECK();
    // the staging array and the scratch are reused by the next call
    MQR_HIP(hipStreamSynchronize(hd->stream));
    return MQR_SUCCESS;
}

}  // namespace mqr
