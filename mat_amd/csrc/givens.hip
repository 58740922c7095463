// QR by Givens rotations (QR(A, Q, R, QR_Givens), QR/QR.h:26-35 with
// QR/__Givens.h:2-21,45-63), on the GPU with the reference's rotation
// sequence and arithmetic, so that Q and R equal the reference's bit for bit
// on inputs without zero pivots.
//
// The reference sweeps column j from the bottom: rotation (i, j) mixes rows
// i-1 and i (i = m-1 .. j+1) with t = R(i,j) / R(i-1,j), c = 1 / sqrt(t t + 1),
// s = t c, then applies the same rotation to columns i-1 and i of Q.
// Rotation (i, j) needs rows i-1 and i after every rotation of the earlier
// columns that touches them, and after (i+1, j).  Giving it the step
// t = (m-1-i) + 2j satisfies all of these, and the rotations that share a
// step touch disjoint row pairs, so:
//   k_givens_r   one workgroup walks the steps; per step it forms (c, s) of
//                the active rotations (one thread each), then applies them to
//                their rows across the columns (threads over columns, R held
//                row-major so a row pair is read coalesced).  Each element
//                sees its rotations in the reference's order, so the rounding
//                is the same.  The (c, s) pairs go to a table for Q.
//   k_givens_q   Q = I, then every row of Q replays all rotations in the
//                reference's order (one thread per row); within a sweep the
//                new column i-1 is carried in a register, so each element of
//                the row is read and written once per sweep.
// Products and sums are rounded one by one (FP contraction is switched off
// in the rotation helpers), as the reference's scalar source.  A zero pivot R(i-1,j) is the one
// deliberate difference: the reference divides by zero and fills R and Q with
// NaN; here the rotation becomes its limit (c = 0, s = sign(R(i,j)); the
// identity when both are zero), which keeps triangular inputs exact.
#include <algorithm>

#include "mqr_internal.h"

namespace mqr {
namespace {

constexpr int GNT = 1024;        // threads of the R workgroup
constexpr int GMAX_ACTIVE = 8192; // rotations one step can hold in LDS

__device__ __forceinline__ double2 rotation(double top, double bot) {
#pragma clang fp contract(off)
    if (top == 0.0) {
        if (bot == 0.0) return make_double2(1.0, 0.0);
        return make_double2(0.0, copysign(1.0, bot));
    }
    const double t = bot / top;
    const double c = 1.0 / sqrt(t * t + 1.0);
    return make_double2(c, t * c);
}

// (a, b) <- (a c + b s, -a s + b c), unfused
__device__ __forceinline__ void rotate(double &a, double &b, double c, double s) {
#pragma clang fp contract(off)
    const double na = a * c + b * s;
    const double nb = -a * s + b * c;
    a = na;
    b = nb;
}

// Rt: m x n row-major (element (r, k) at Rt[r * n + k]); CS: (c, s) of
// rotation (i, j) at CS[j * m + i].
__global__ __launch_bounds__(GNT) void k_givens_r(int m, int n, double *__restrict__ Rt,
                                                  double2 *__restrict__ CS) {
    __shared__ double2 cs[GMAX_ACTIVE];
    const int tid = threadIdx.x;
    const int jmax = min(n, m - 1) - 1;  // last column with a rotation
    if (jmax < 0) return;
    const int steps = m - 1 + jmax;      // last rotation (j+1, jmax) at step m-2+jmax
    for (int t = 0; t < steps; ++t) {
        const int jlo = max(0, t - m + 2), jhi = min(jmax, t >> 1);
        const int cnt = jhi - jlo + 1;
        for (int a = tid; a < cnt; a += GNT) {
            const int j = jlo + a, i = m - 1 - t + 2 * j;
            const double2 r = rotation(Rt[(int64_t)(i - 1) * n + j], Rt[(int64_t)i * n + j]);
            cs[a] = r;
            CS[(int64_t)j * m + i] = r;
        }
        __syncthreads();
        for (int a = 0; a < cnt; ++a) {
            const int j = jlo + a, i = m - 1 - t + 2 * j;
            const double2 r = cs[a];
            double *top = Rt + (int64_t)(i - 1) * n, *bot = Rt + (int64_t)i * n;
            for (int k = j + tid; k < n; k += GNT) {
                double x = top[k], y = bot[k];
                rotate(x, y, r.x, r.y);
                top[k] = x;
                bot[k] = y;
            }
        }
        __syncthreads();
    }
}

// Q (m x m, ldq) = I with the rotations of CS applied on the right in the
// reference's order; thread = row.
__global__ __launch_bounds__(64) void k_givens_q(int m, int n, const double2 *__restrict__ CS,
                                                 double *__restrict__ Q, int64_t ldq) {
    const int r = blockIdx.x * 64 + threadIdx.x;
    if (r >= m) return;
    for (int c = 0; c < m; ++c) Q[(int64_t)c * ldq + r] = r == c ? 1.0 : 0.0;
    const int jmax = min(n, m - 1) - 1;
    for (int j = 0; j <= jmax; ++j) {
        const double2 *csj = CS + (int64_t)j * m;
        double carry = Q[(int64_t)(m - 1) * ldq + r];  // column i of the next rotation
        for (int i = m - 1; i > j; --i) {
            double a = Q[(int64_t)(i - 1) * ldq + r];
            const double2 g = csj[i];
            rotate(a, carry, g.x, g.y);
            Q[(int64_t)i * ldq + r] = carry;  // column i is final for this sweep
            carry = a;
        }
        Q[(int64_t)j * ldq + r] = carry;
    }
}

// Rt (m x n row-major) <-> R (column-major, ld)
__global__ void k_to_rowmajor(int m, int n, const double *__restrict__ A, int64_t lda,
                              double *__restrict__ Rt) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (int64_t)m * n) return;
    const int r = (int)(e / n), k = (int)(e % n);
    Rt[e] = A[(int64_t)k * lda + r];
}

__global__ void k_from_rowmajor(int m, int n, const double *__restrict__ Rt, double *__restrict__ R,
                                int64_t ldr) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (int64_t)m * n) return;
    const int r = (int)(e % m), k = (int)(e / m);
    R[(int64_t)k * ldr + r] = Rt[(int64_t)r * n + k];
}

}  // namespace

// Workspace (bytes) of qr_givens_f64: R row-major and the (c, s) table.
size_t givens_ws_bytes(int64_t m, int64_t n) {
    return sizeof(double) * (size_t)(m * n) + sizeof(double2) * (size_t)(m * std::min(m, n)) + 512;
}

// Largest number of rotations one step runs (the LDS table of k_givens_r).
bool givens_supported(int64_t m, int64_t n) { return std::min(n, m / 2 + 1) <= GMAX_ACTIVE; }

int qr_givens_f64(mqr_handle_t hd, int64_t m, int64_t n, const double *A, int64_t lda, double *Q,
                  int64_t ldq, double *R, int64_t ldr) {
    MQR_CHECK(ws_reserve(hd, hd->ws, givens_ws_bytes(m, n)));
    double *Rt = (double *)hd->ws.ptr;
    double2 *CS = (double2 *)(((uintptr_t)(Rt + m * n) + 255) & ~(uintptr_t)255);
    hipStream_t s = hd->stream;
    const int64_t mn = m * n;
    if (mn > 0) {
        hipLaunchKernelGGL(k_to_rowmajor, dim3((unsigned)((mn + 255) / 256)), dim3(256), 0, s,
                           (int)m, (int)n, A, lda, Rt);
        MQR_LAUNCH_CHECK();
        ProfScope ps(K_PANEL_SMALL, 6.0 * (double)m * n * n, 32.0 * (double)m * n * n, s);
        hipLaunchKernelGGL(k_givens_r, dim3(1), dim3(GNT), 0, s, (int)m, (int)n, Rt, CS);
        MQR_LAUNCH_CHECK();
    }
    {
        ProfScope ps(K_APPLYQ, 6.0 * (double)m * m * n, 16.0 * (double)m * m * n, s);
        hipLaunchKernelGGL(k_givens_q, dim3((unsigned)((m + 63) / 64)), dim3(64), 0, s, (int)m,
                           (int)n, (const double2 *)CS, Q, ldq);
        MQR_LAUNCH_CHECK();
    }
    if (mn > 0) {
        hipLaunchKernelGGL(k_from_rowmajor, dim3((unsigned)((mn + 255) / 256)), dim3(256), 0, s,
                           (int)m, (int)n, (const double *)Rt, R, ldr);
        MQR_LAUNCH_CHECK();
    }
    return MQR_SUCCESS;
}

}  // namespace mqr
