/*
 * mqr_oracle.c -- TEST INFRASTRUCTURE, NOT PRODUCT CODE.
 *
 * A from-scratch CPU restatement of Mat-8.3.1's Householder QR hot path,
 * written to serve as the parity oracle for the HIP implementation and as the
 * timed CPU baseline.  Only tests/, the smoke check and bench.py's
 * cpu_baseline leg use it.  Every function names the reference lines whose
 * arithmetic it follows; the operation order (k-order fused multiply-subtract
 * in the MR x NR tile, four-way split dots, axpy as fma) is kept so that the
 * results sit within a few ulps of the reference's own.
 *
 * Build-time block sizes (Mat.h:57-109) come from ORC_ISA_AVX512 or the AVX2
 * default; see the Makefile next to this file.
 */
#include "mqr_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#if defined(ORC_ISA_AVX512)
#define MR 16
#define NR 12
#define MC 1152
#define KC 72
#define NC 144
#define BS 48
#else
#define MR 8
#define NR 6
#define MC 576
#define KC 144
#define NC 144
#define BS 96
#endif

static inline int imin(int a, int b) { return a < b ? a : b; }
static inline int imax(int a, int b) { return a > b ? a : b; }
static inline size_t up8(size_t x) { return (x + 7) & ~(size_t)7; }

void orc_config(int *out) {
    out[0] = MR; out[1] = NR; out[2] = MC; out[3] = KC; out[4] = NC; out[5] = BS;
}

/* ------------------------------------------------------------------ */
/* Level-1 helpers: Calculate/Calculate.h:45-76                        */
/* ------------------------------------------------------------------ */

/* Four partial sums over groups of four, summed pairwise-left, then the
 * tail added one by one (Calculate.h:61-76). */
double orc_dot(const double *a, const double *b, int len) {
    double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
    const int n4 = len & ~3;
    for (int i = 0; i < n4; i += 4) {
        s0 = fma(a[i], b[i], s0);
        s1 = fma(a[i + 1], b[i + 1], s1);
        s2 = fma(a[i + 2], b[i + 2], s2);
        s3 = fma(a[i + 3], b[i + 3], s3);
    }
    double r = s0 + s1 + s2 + s3;
    for (int i = n4; i < len; ++i) r = fma(a[i], b[i], r);
    return r;
}

/* ans = a + b * q, fused (the AVX path's fmadd, Calculate/__AVX2_double.h:101-118). */
static void axpy_into(double *ans, const double *a, const double *b, double q, int len) {
    for (int i = 0; i < len; ++i) ans[i] = fma(b[i], q, a[i]);
}

/* ------------------------------------------------------------------ */
/* Packing: Mul/__pack.h:2-258 (layouts), with this build's MR / NR     */
/* ------------------------------------------------------------------ */

/* MR-row panels of an mc x kc block; panel ip holds element (i, p) at
 * ip*kc*MR + p*MR + i, zero padded (__pack.h:2-22). */
void orc_pack_A(const double *A, int64_t ldA, int mc, int kc, double *Ap) {
    for (int ip = 0; ip * MR < mc; ++ip) {
        const int mm = imin(MR, mc - ip * MR);
        double *dst = Ap + (size_t)ip * kc * MR;
        for (int p = 0; p < kc; ++p) {
            const double *src = A + (size_t)p * ldA + (size_t)ip * MR;
            int i = 0;
            for (; i < mm; ++i) dst[i] = src[i];
            for (; i < MR; ++i) dst[i] = 0.0;
            dst += MR;
        }
    }
}

/* Panels of V^T: rows of V^T are columns of V; element (i, q) of panel ip is
 * V[pc + q, ip*MR + i] (__pack.h:23-43). */
static void pack_AT(const double *V, int64_t ldV, int pc, int mc, int kc, double *Ap) {
    for (int ip = 0; ip * MR < mc; ++ip) {
        const int mm = imin(MR, mc - ip * MR);
        double *dst = Ap + (size_t)ip * kc * MR;
        for (int q = 0; q < kc; ++q) {
            const double *src = V + (size_t)(ip * MR) * ldV + pc + q;
            int i = 0;
            for (; i < mm; ++i) dst[i] = src[(size_t)i * ldV];
            for (; i < MR; ++i) dst[i] = 0.0;
            dst += MR;
        }
    }
}

/* NR-column panels of a kc x nc block; element (p, j) of panel jp at
 * jp*kc*NR + p*NR + j (__pack.h:45-63). */
void orc_pack_B(const double *B, int64_t ldB, int kc, int nc, double *Bp) {
    for (int jp = 0; jp * NR < nc; ++jp) {
        const int nn = imin(NR, nc - jp * NR);
        double *dst = Bp + (size_t)jp * kc * NR;
        for (int p = 0; p < kc; ++p) {
            int j = 0;
            for (; j < nn; ++j) dst[(size_t)p * NR + j] = B[(size_t)(jp * NR + j) * ldB + p];
            for (; j < NR; ++j) dst[(size_t)p * NR + j] = 0.0;
        }
    }
}

/* V (rows x cols, unit lower trapezoidal, reflectors below the diagonal) as
 * A-panels: the first ceil(cols/MR)*MR rows carry the implicit unit diagonal
 * and zeros above it; panel r only holds k-columns p < r + m
 * (__pack.h:108-149). */
static void pack_L(const double *V, int64_t ldV, int rows, int cols, double *Lp) {
    const int s = imin(rows, ((cols + MR - 1) / MR) * MR);
    for (int r = 0; r < s; r += MR) {
        const int m = imin(MR, s - r);
        const int pmax = imin(cols, r + m);
        double *dst = Lp + (size_t)r * cols;
        for (int p = 0; p < pmax; ++p) {
            const double *Vc = V + (size_t)p * ldV;
            int i = 0;
            for (; i < m; ++i) {
                const int row = r + i;
                dst[i] = row > p ? Vc[row] : (row == p ? 1.0 : 0.0);
            }
            for (; i < MR; ++i) dst[i] = 0.0;
            dst += MR;
        }
    }
    if (s < rows) orc_pack_A(V + s, ldV, rows - s, cols, Lp + (size_t)(s / MR) * cols * MR);
}

/* V^T as A-panels: the cols x cols head is unit upper, then KC-row chunks of
 * the rest of V^T (__pack.h:151-189 and QR/__Householder_Blocked.h:121-141). */
static void pack_VT(const double *V, int64_t ldV, int rows, int cols, double *VTp) {
    for (int r = 0; r < cols; r += MR) {
        const int m = imin(MR, cols - r);
        double *dst = VTp + (size_t)r * cols;
        for (int p = 0; p < cols; ++p) {
            int i = 0;
            for (; i < m; ++i) {
                const int row = r + i;
                dst[i] = row < p ? V[(size_t)row * ldV + p] : (row == p ? 1.0 : 0.0);
            }
            for (; i < MR; ++i) dst[i] = 0.0;
            dst += MR;
        }
    }
    const int N = (cols + MR - 1) / MR;
    double *dst = VTp + (size_t)N * cols * MR;
    for (int r = cols; r < rows; r += KC) {
        const int kc = imin(KC, rows - r);
        pack_AT(V, ldV, r, cols, kc, dst);
        dst += (size_t)N * KC * MR;
    }
}

/* Upper-triangular T (k x k, ld = k) as A-panels for the upper-triangular
 * driver: the panel of rows r starts at r*(k+MR) and holds k-columns p >= r
 * (__pack.h:191-218). */
static void pack_U(const double *T, int k, double *Up) {
    for (int r = 0; r < k; r += MR) {
        const int m = imin(MR, k - r);
        double *dst = Up + (size_t)r * (k + MR);
        for (int p = r; p < k; ++p) {
            int i = 0;
            for (; i < m; ++i) {
                const int row = r + i;
                dst[i] = row > p ? 0.0 : T[(size_t)p * k + row];
            }
            for (; i < MR; ++i) dst[i] = 0.0;
            dst += MR;
        }
    }
}

/* T^T (lower triangular) as A-panels for the lower-triangular driver
 * (__pack.h:220-246). */
static void pack_UT(const double *T, int k, double *UTp) {
    for (int r = 0; r < k; r += MR) {
        const int m = imin(MR, k - r);
        const int pmax = imin(k, r + m);
        double *dst = UTp + (size_t)r * k;
        for (int p = 0; p < pmax; ++p) {
            int i = 0;
            for (; i < m; ++i) {
                const int row = r + i;
                dst[i] = row < p ? 0.0 : T[(size_t)row * k + p];
            }
            for (; i < MR; ++i) dst[i] = 0.0;
            dst += MR;
        }
    }
}

/* ------------------------------------------------------------------ */
/* MR x NR micro-kernel: Mul/__micro_kernel_packed_double.h:763-924    */
/* (AVX2 8x6) and :301-585 (AVX-512 16x12).  C is loaded into the       */
/* accumulators, updated in k order with one fused multiply-add (add)   */
/* or multiply-subtract (nmul) per element and step, then stored.       */
/* ------------------------------------------------------------------ */
#if defined(ORC_SIMD)
/* The same micro-kernel with the accumulators in vector registers (the
 * ORC_SIMD builds, used as the timed CPU baseline): C in MR/W vectors per
 * column, per k step one vector load of the packed A column piece and one
 * broadcast of each packed B value, one fused multiply-add (or negated
 * multiply-add) per element and step in k order, as in the scalar loop.
 * The rounding is therefore identical element by element; partial edge
 * tiles take the scalar loop. */
#include <immintrin.h>
#if defined(ORC_ISA_AVX512)
#define VW 8
typedef __m512d vd;
#define V_LOAD(p) _mm512_loadu_pd(p)
#define V_STORE(p, v) _mm512_storeu_pd((p), (v))
#define V_BCAST(x) _mm512_set1_pd(x)
#define V_FMADD(a, b, c) _mm512_fmadd_pd((a), (b), (c))
#define V_FNMADD(a, b, c) _mm512_fnmadd_pd((a), (b), (c))
#else
#define VW 4
typedef __m256d vd;
#define V_LOAD(p) _mm256_loadu_pd(p)
#define V_STORE(p, v) _mm256_storeu_pd((p), (v))
#define V_BCAST(x) _mm256_set1_pd(x)
#define V_FMADD(a, b, c) _mm256_fmadd_pd((a), (b), (c))
#define V_FNMADD(a, b, c) _mm256_fnmadd_pd((a), (b), (c))
#endif
#define MV (MR / VW)

static void micro_kernel_simd(const double *Ap, const double *Bp, double *C, int64_t ldC, int k,
                              int add) {
    vd acc[NR][MV];
    for (int j = 0; j < NR; ++j)
        for (int v = 0; v < MV; ++v) acc[j][v] = V_LOAD(C + (size_t)j * ldC + v * VW);
    if (add) {
        for (int p = 0; p < k; ++p) {
            const double *a = Ap + (size_t)p * MR;
            const double *b = Bp + (size_t)p * NR;
            vd av[MV];
            for (int v = 0; v < MV; ++v) av[v] = V_LOAD(a + v * VW);
            for (int j = 0; j < NR; ++j) {
                const vd bj = V_BCAST(b[j]);
                for (int v = 0; v < MV; ++v) acc[j][v] = V_FMADD(av[v], bj, acc[j][v]);
            }
        }
    } else {
        for (int p = 0; p < k; ++p) {
            const double *a = Ap + (size_t)p * MR;
            const double *b = Bp + (size_t)p * NR;
            vd av[MV];
            for (int v = 0; v < MV; ++v) av[v] = V_LOAD(a + v * VW);
            for (int j = 0; j < NR; ++j) {
                const vd bj = V_BCAST(b[j]);
                for (int v = 0; v < MV; ++v) acc[j][v] = V_FNMADD(av[v], bj, acc[j][v]);
            }
        }
    }
    for (int j = 0; j < NR; ++j)
        for (int v = 0; v < MV; ++v) V_STORE(C + (size_t)j * ldC + v * VW, acc[j][v]);
}
#endif

void orc_micro_kernel(const double *Ap, const double *Bp, double *C, int64_t ldC,
                      int k, int mm, int nn, int add) {
#if defined(ORC_SIMD)
    if (mm == MR && nn == NR) {
        micro_kernel_simd(Ap, Bp, C, ldC, k, add);
        return;
    }
#endif
    double acc[NR][MR];
    const int full = (mm == MR && nn == NR);
    for (int j = 0; j < NR; ++j)
        for (int i = 0; i < MR; ++i)
            acc[j][i] = (full || (i < mm && j < nn)) ? C[(size_t)j * ldC + i] : 0.0;
    if (add) {
        for (int p = 0; p < k; ++p) {
            const double *a = Ap + (size_t)p * MR;
            const double *b = Bp + (size_t)p * NR;
            for (int j = 0; j < NR; ++j) {
                const double bj = b[j];
                for (int i = 0; i < MR; ++i) acc[j][i] = fma(a[i], bj, acc[j][i]);
            }
        }
    } else {
        for (int p = 0; p < k; ++p) {
            const double *a = Ap + (size_t)p * MR;
            const double *b = Bp + (size_t)p * NR;
            for (int j = 0; j < NR; ++j) {
                const double bj = b[j];
                for (int i = 0; i < MR; ++i) acc[j][i] = fma(-a[i], bj, acc[j][i]);
            }
        }
    }
    for (int j = 0; j < NR; ++j)
        for (int i = 0; i < MR; ++i)
            if (full || (i < mm && j < nn)) C[(size_t)j * ldC + i] = acc[j][i];
}

/* ------------------------------------------------------------------ */
/* Macro-kernel drivers: Mul/Mul.h:123-152, :239-322                   */
/* ------------------------------------------------------------------ */
static void mulAB_packed(const double *Ap, const double *Bp, double *C, int64_t ldC,
                         int mc, int nc, int kc) {
    const int mp = (mc + MR - 1) / MR, np = (nc + NR - 1) / NR;
    for (int jp = 0; jp < np; ++jp) {
        const int nn = imin(NR, nc - jp * NR);
        for (int ip = 0; ip < mp; ++ip) {
            const int mm = imin(MR, mc - ip * MR);
            orc_micro_kernel(Ap + (size_t)ip * kc * MR, Bp + (size_t)jp * kc * NR,
                             C + (size_t)(jp * NR) * ldC + ip * MR, ldC, kc, mm, nn, 1);
        }
    }
}

void orc_mulAB_packed_n(const double *Ap, const double *Bp, double *C, int64_t ldC,
                        int m, int n, int k) {
    const int mp = (m + MR - 1) / MR, np = (n + NR - 1) / NR;
    for (int jp = 0; jp < np; ++jp) {
        const int nn = imin(NR, n - jp * NR);
        for (int ip = 0; ip < mp; ++ip) {
            const int mm = imin(MR, m - ip * MR);
            orc_micro_kernel(Ap + (size_t)ip * k * MR, Bp + (size_t)jp * k * NR,
                             C + (size_t)(jp * NR) * ldC + ip * MR, ldC, k, mm, nn, 0);
        }
    }
}

/* Lower-triangular A: the tile of rows ip*MR.. only needs k < ip*MR + mm. */
static void mulAB_lower_tri_packed_n(const double *Ap, const double *Bp, double *C, int64_t ldC,
                                     int m, int n, int k) {
    const int mp = (m + MR - 1) / MR, np = (n + NR - 1) / NR;
    for (int jp = 0; jp < np; ++jp) {
        const int nn = imin(NR, n - jp * NR);
        for (int ip = 0; ip < mp; ++ip) {
            const int mm = imin(MR, m - ip * MR);
            const int keff = imin(k, ip * MR + mm);
            if (keff <= 0) continue;
            orc_micro_kernel(Ap + (size_t)ip * k * MR, Bp + (size_t)jp * k * NR,
                             C + (size_t)(jp * NR) * ldC + ip * MR, ldC, keff, mm, nn, 0);
        }
    }
}

/* Upper-triangular A: the tile of rows ip*MR.. starts at k = ip*MR. */
static void mulAB_upper_tri_packed_n(const double *Ap, const double *Bp, double *C, int64_t ldC,
                                     int m, int n, int k) {
    const int mp = (m + MR - 1) / MR, np = (n + NR - 1) / NR;
    for (int jp = 0; jp < np; ++jp) {
        const int nn = imin(NR, n - jp * NR);
        for (int ip = 0; ip < mp; ++ip) {
            const int mm = imin(MR, m - ip * MR);
            const int k0 = imax(0, ip * MR);
            const int keff = k - k0;
            if (keff <= 0) continue;
            orc_micro_kernel(Ap + (size_t)ip * k * MR + (size_t)k0 * MR,
                             Bp + (size_t)jp * k * NR + (size_t)k0 * NR,
                             C + (size_t)(jp * NR) * ldC + ip * MR, ldC, keff, mm, nn, 0);
        }
    }
}

/* C = A * B through the blocked, packed GEMM (Mul/Mul.h:154-202). */
void orc_gemm(int M, int N, int K, const double *A, int64_t ldA,
              const double *B, int64_t ldB, double *C, int64_t ldC) {
    for (int j = 0; j < N; ++j) memset(C + (size_t)j * ldC, 0, sizeof(double) * (size_t)M);
    double *Ap = (double *)malloc(sizeof(double) * ((size_t)(MC + MR - 1) / MR) * KC * MR);
    double *Bp = (double *)malloc(sizeof(double) * ((size_t)(NC + NR - 1) / NR) * KC * NR);
    for (int pc = 0; pc < K; pc += KC) {
        const int kc = imin(KC, K - pc);
        for (int ic = 0; ic < M; ic += MC) {
            const int mc = imin(MC, M - ic);
            orc_pack_A(A + (size_t)pc * ldA + ic, ldA, mc, kc, Ap);
            for (int jc = 0; jc < N; jc += NC) {
                const int nc = imin(NC, N - jc);
                orc_pack_B(B + pc + (size_t)jc * ldB, ldB, kc, nc, Bp);
                mulAB_packed(Ap, Bp, C + (size_t)jc * ldC + ic, ldC, mc, nc, kc);
            }
        }
    }
    free(Ap);
    free(Bp);
}

/* ------------------------------------------------------------------ */
/* Workspace arena: QR/__Householder_Blocked.h:11-55                   */
/* ------------------------------------------------------------------ */
typedef struct { double *packA, *packB, *W, *W2, *buf; } arena_t;

static size_t work1_doubles(int m) {
    const size_t v_pack = ((size_t)(m + MR - 1) / MR) * BS * MR;
    const size_t npb = ((size_t)BS + MR - 1) / MR;
    const size_t nfull = ((size_t)m + KC - 1) / KC;
    const size_t vt_pack = npb * BS * MR + npb * nfull * KC * MR;
    const size_t u_pack = npb * MR * (BS + MR); /* pack_U reaches r*(k+MR) */
    size_t w1 = v_pack > vt_pack ? v_pack : vt_pack;
    return up8(w1 > u_pack ? w1 : u_pack);
}
static size_t work2_doubles(void) {
    const size_t r_pack = ((size_t)(NC + NR - 1) / NR) * KC * NR;
    const size_t w_pack = ((size_t)(NC + NR - 1) / NR) * BS * NR;
    return up8(r_pack > w_pack ? r_pack : w_pack);
}

size_t orc_work_doubles(int m, int n) {
    const int k = imin(m, n), nW = imax(m, n);
    return work1_doubles(m) + work2_doubles() + 2 * up8((size_t)BS * nW) +
           up8((size_t)m + 2 * BS + k);
}

static arena_t carve(double *work, int m, int n) {
    arena_t a;
    const int nW = imax(m, n);
    a.packA = work;
    a.packB = a.packA + work1_doubles(m);
    a.W = a.packB + work2_doubles();
    a.W2 = a.W + up8((size_t)BS * nW);
    a.buf = a.W2 + up8((size_t)BS * nW);
    return a;
}

/* ------------------------------------------------------------------ */
/* Panel: QR/__Householder_Blocked.h:62-84                             */
/* LAPACK dlarfg convention without its rescaling: cr = sign(w0)*||w||  */
/* (sign(0) = +), R_rr = -cr, v = w[1:]/(w0+cr), tau = (w0+cr)/cr; a    */
/* zero column gets tau = 0 and is skipped.                              */
/* ------------------------------------------------------------------ */
void orc_panel(int h, int b, double *P, int64_t ldP, double *tau, double *v) {
    for (int r = 0; r < b; ++r) {
        double *w = P + (size_t)r * ldP + r;
        const int len = h - r;
        double cr = sqrt(orc_dot(w, w, len));
        if (cr == 0.0) { tau[r] = 0.0; continue; }
        if (w[0] < 0.0) cr = -cr;
        const double w0 = w[0] + cr;
        const double hr = 1.0 / w0;
        w[0] = -cr;
        for (int i = 1; i < len; ++i) w[i] = w[i] * hr;
        v[0] = 1.0;
        for (int i = 1; i < len; ++i) v[i] = w[i];
        tau[r] = w0 / cr;
        for (int c = r + 1; c < b; ++c) {
            double *col = P + (size_t)c * ldP + r;
            const double d = col[0] + orc_dot(v + 1, col + 1, len - 1);
            axpy_into(col, col, v, -d * tau[r], len);
        }
    }
}

/* ------------------------------------------------------------------ */
/* Compact-WY T, forward columnwise: QR/__Householder_Blocked.h:86-103 */
/* ------------------------------------------------------------------ */
void orc_larft(int h, int b, const double *V, int64_t ldV, const double *tau,
               double *T, double *s, double *w) {
    const int ldT = b;
    T[0] = tau[0];
    for (int i = 1; i < b; ++i) {
        const double *Vi = V + (size_t)i * ldV;
        for (int l = 0; l < i; ++l) {
            const double *Vl = V + (size_t)l * ldV;
            s[l] = Vl[i] + orc_dot(Vl + i + 1, Vi + i + 1, h - i - 1);
        }
        for (int r = 0; r < i; ++r) w[r] = 0.0;
        for (int c = 0; c < i; ++c) axpy_into(w, w, T + (size_t)c * ldT, s[c], c + 1);
        for (int r = 0; r < i; ++r) T[(size_t)i * ldT + r] = w[r] * -tau[i];
        T[(size_t)i * ldT + i] = tau[i];
    }
}

/* ------------------------------------------------------------------ */
/* Block reflector application, QR/__Householder_Blocked.h:143-212      */
/* (transposed, factorisation) and :271-331 (plain, Q formation):       */
/*   W  = -V^T C     (unit-upper head of V^T, then KC chunks)            */
/*   W2 = -T^T W     (transposed) or -T W (plain)                        */
/*   C -= V W2       (unit-lower head of V, then MC chunks)              */
/* processed NC columns of C at a time.                                  */
/* ------------------------------------------------------------------ */
static void larfb(int h, int k, int ncols, const double *V, int64_t ldV, const double *T,
                  double *C, int64_t ldC, double *packA, double *packB, double *W, double *W2,
                  int transposed) {
    const int N = (k + MR - 1) / MR;
    const size_t tri = (size_t)N * k * MR;
    const size_t fullblk = (size_t)N * KC * MR;

    pack_VT(V, ldV, h, k, packA);
    for (int j = 0; j < ncols; j += NC) {
        const int nc = imin(NC, ncols - j);
        double *Wj = W + (size_t)j * k;
        memset(Wj, 0, sizeof(double) * (size_t)k * nc);
        double *Cj = C + (size_t)j * ldC;
        orc_pack_B(Cj, ldC, k, nc, packB);
        mulAB_upper_tri_packed_n(packA, packB, Wj, k, k, nc, k);
        const double *TA = packA + tri;
        for (int r = k; r < h; r += KC) {
            const int kc = imin(KC, h - r);
            orc_pack_B(Cj + r, ldC, kc, nc, packB);
            orc_mulAB_packed_n(TA, packB, Wj, k, k, nc, kc);
            TA += fullblk;
        }
    }

    if (transposed) pack_UT(T, k, packA); else pack_U(T, k, packA);
    for (int j = 0; j < ncols; j += NC) {
        const int nc = imin(NC, ncols - j);
        double *Wj = W + (size_t)j * k, *W2j = W2 + (size_t)j * k;
        orc_pack_B(Wj, k, k, nc, packB);
        memset(W2j, 0, sizeof(double) * (size_t)k * nc);
        if (transposed) mulAB_lower_tri_packed_n(packA, packB, W2j, k, k, nc, k);
        else mulAB_upper_tri_packed_n(packA, packB, W2j, k, k, nc, k);
    }

    pack_L(V, ldV, h, k, packA);
    const int tri_n = imin(h, ((k + MR - 1) / MR) * MR);
    for (int j = 0; j < ncols; j += NC) {
        const int nc = imin(NC, ncols - j);
        const double *W2j = W2 + (size_t)j * k;
        orc_pack_B(W2j, k, k, nc, packB);
        double *Cj = C + (size_t)j * ldC;
        mulAB_lower_tri_packed_n(packA, packB, Cj, ldC, tri_n, nc, k);
        for (int i = tri_n; i < h; i += MC) {
            const int mc = imin(MC, h - i);
            orc_mulAB_packed_n(packA + (size_t)(i / MR) * k * MR, packB, Cj + i, ldC, mc, nc, k);
        }
    }
}

void orc_larfb_t(int h, int b, int nc, const double *V, int64_t ldV,
                 const double *T, double *C, int64_t ldC, double *work) {
    arena_t a = carve(work, h, imax(h, nc));
    larfb(h, b, nc, V, ldV, T, C, ldC, a.packA, a.packB, a.W, a.W2, 1);
}

void orc_larfb_n(int h, int b, int nc, const double *V, int64_t ldV,
                 const double *T, double *C, int64_t ldC, double *work) {
    arena_t a = carve(work, h, imax(h, nc));
    larfb(h, b, nc, V, ldV, T, C, ldC, a.packA, a.packB, a.W, a.W2, 0);
}

/* ------------------------------------------------------------------ */
/* Factorisation driver: QR/__Householder_Blocked.h:214-245             */
/* ------------------------------------------------------------------ */
void orc_geqrf_blocked(int m, int n, double *R, double *T_all, double *tau, double *work) {
    const int k = imin(m, n);
    arena_t a = carve(work, m, n);
    double *v = a.buf, *s = v + m, *w = s + BS;
    for (int j = 0; j < k; j += BS) {
        const int b = imin(BS, k - j);
        const int h = m - j;
        double *V = R + (size_t)j * m + j;
        orc_panel(h, b, V, m, tau + j, v);
        orc_larft(h, b, V, m, tau + j, T_all + (size_t)j * BS, s, w);
        if (j + b < n)
            larfb(h, b, n - j - b, V, m, T_all + (size_t)j * BS,
                  R + (size_t)(j + b) * m + j, m, a.packA, a.packB, a.W, a.W2, 1);
    }
}

/* ------------------------------------------------------------------ */
/* Explicit full Q: QR/__Householder_Blocked.h:247-333.  Q = I, then    */
/* blocks from last to first applied to all m columns of Q[j:m, :].     */
/* ------------------------------------------------------------------ */
void orc_orgqr_full(int m, int n, const double *VR, const double *T_all, double *Q, double *work) {
    const int k = imin(m, n);
    arena_t a = carve(work, m, n);
    for (int c = 0; c < m; ++c) {
        memset(Q + (size_t)c * m, 0, sizeof(double) * (size_t)m);
        Q[(size_t)c * m + c] = 1.0;
    }
    if (k <= 0) return;
    const int last = ((k - 1) / BS) * BS;
    for (int j = last; j >= 0; j -= BS) {
        const int b = imin(BS, k - j);
        larfb(m - j, b, m, VR + (size_t)j * m + j, m, T_all + (size_t)j * BS,
              Q + j, m, a.packA, a.packB, a.W, a.W2, 0);
    }
}

/* ------------------------------------------------------------------ */
/* Unblocked branch: QR/__HouseHolder.h:178-208 (reflectors normalised  */
/* so that H = I - w w^T) and :47-79 (Q built backward).                 */
/* ------------------------------------------------------------------ */
void orc_qr_unblocked(int m, int n, double *Q, double *R) {
    const int k = imin(m, n);
    /* reflector r lives in its own column slot of Wst (length m - r) */
    double *Wst = (double *)calloc((size_t)m * (k > 0 ? k : 1), sizeof(double));
    for (int r = 0; r < k; ++r) {
        double *w = Wst + (size_t)r * m;
        const int len = m - r;
        memcpy(w, R + (size_t)r * m + r, sizeof(double) * (size_t)len);
        double cr = sqrt(orc_dot(w, w, len));
        if (cr == 0.0) continue;
        if (w[0] < 0.0) cr = -cr;
        w[0] += cr;
        const double hr = 1.0 / sqrt(cr * w[0]);
        for (int i = 0; i < len; ++i) w[i] = w[i] * hr;
        /* __ApplyHouseHolderOnTheLeft2 (:39-44): every column from r on */
        for (int c = r; c < n; ++c) {
            double *col = R + (size_t)c * m + r;
            axpy_into(col, col, w, -orc_dot(w, col, len), len);
        }
    }
    /* __HouseHolder_Q1: Q = I; for r = k-1..0 the trailing block
     * Q[r:, r:] = (I - w w^T) diag(1, Q[r+1:, r+1:]). */
    for (int c = 0; c < m; ++c) {
        memset(Q + (size_t)c * m, 0, sizeof(double) * (size_t)m);
        Q[(size_t)c * m + c] = 1.0;
    }
    for (int r = k - 1; r >= 0; --r) {
        const double *w = Wst + (size_t)r * m;
        const int len = m - r;
        for (int i = 1; i < len; ++i) {
            double *col = Q + (size_t)(r + i) * m + r; /* column r+i, rows r.. */
            const double t = -orc_dot(w + 1, col + 1, len - 1);
            col[0] = w[0] * t;
            axpy_into(col + 1, col + 1, w + 1, t, len - 1);
        }
        double *c0 = Q + (size_t)r * m + r;
        const double w0 = w[0];
        for (int i = 0; i < len; ++i) c0[i] = w[i] * -w0;
        c0[0] += 1.0;
    }
    free(Wst);
}

/* ------------------------------------------------------------------ */
/* QR(A, Q, R, QR_Householder): QR/QR.h:7-24 plus the blocked wrapper's  */
/* R clean-up (QR/__Householder_Blocked.h:335-353).                       */
/* ------------------------------------------------------------------ */
int orc_qr_householder(int m, int n, const double *A, double *Q, double *R) {
    memcpy(R, A, sizeof(double) * (size_t)m * n);
    if (imin(m, n) <= 2 * BS) {
        orc_qr_unblocked(m, n, Q, R);
        return 0;
    }
    const int k = imin(m, n);
    double *T_all = (double *)calloc((size_t)((k + BS - 1) / BS) * BS * BS, sizeof(double));
    double *tau = (double *)calloc((size_t)k, sizeof(double));
    double *work = (double *)calloc(orc_work_doubles(m, n), sizeof(double));
    orc_geqrf_blocked(m, n, R, T_all, tau, work);
    orc_orgqr_full(m, n, R, T_all, Q, work);
    for (int j = 0; j < n; ++j) {
        const int below = m - j - 1;
        if (below <= 0) break;
        memset(R + (size_t)j * m + j + 1, 0, sizeof(double) * (size_t)below);
    }
    free(T_all);
    free(tau);
    free(work);
    return 1;
}

/* ------------------------------------------------------------------ */
/* Least squares (no reference counterpart)                             */
/* ------------------------------------------------------------------ */
void orc_trsv_upper(int n, const double *R, int64_t ldR, double *x) {
    for (int j = n - 1; j >= 0; --j) {
        x[j] = x[j] / R[(size_t)j * ldR + j];
        const double xj = -x[j];
        const double *col = R + (size_t)j * ldR;
        for (int i = 0; i < j; ++i) x[i] = fma(col[i], xj, x[i]);
    }
}

void orc_gels(int m, int n, const double *A, const double *b, double *x) {
    const int k = imin(m, n);
    double *R = (double *)malloc(sizeof(double) * (size_t)m * n);
    double *c = (double *)malloc(sizeof(double) * (size_t)m);
    double *T_all = (double *)calloc((size_t)((k + BS - 1) / BS) * BS * BS, sizeof(double));
    double *tau = (double *)calloc((size_t)(k > 0 ? k : 1), sizeof(double));
    double *work = (double *)calloc(orc_work_doubles(m, n), sizeof(double));
    memcpy(R, A, sizeof(double) * (size_t)m * n);
    memcpy(c, b, sizeof(double) * (size_t)m);
    orc_geqrf_blocked(m, n, R, T_all, tau, work);
    arena_t a = carve(work, m, n);
    for (int j = 0; j < k; j += BS) {
        const int bb = imin(BS, k - j);
        larfb(m - j, bb, 1, R + (size_t)j * m + j, m, T_all + (size_t)j * BS, c + j, m,
              a.packA, a.packB, a.W, a.W2, 1);
    }
    memcpy(x, c, sizeof(double) * (size_t)n);
    orc_trsv_upper(n, R, m, x);
    free(R); free(c); free(T_all); free(tau); free(work);
}

/* ------------------------------------------------------------------ */
/* QR(A, Q, R, QR_Givens): QR/QR.h:26-35 with QR/__Givens.h:2-21,45-63 */
/* ------------------------------------------------------------------ */
/* R = A, Q = I_m; for each column j and i = m-1 down to j+1 a rotation of
 * rows (i-1, i) built from t = R(i,j) / R(i-1,j): c = 1 / sqrt(t t + 1),
 * s = t c (:50-55); rows i-1, i of R from column j on become
 * (a c + b s, -a s + b c) (__ApplyGivensOnTheLeft, :13-23) and columns
 * i-1, i of Q the same (__ApplyGivensOnTheRight, :2-11).  The source's
 * plain products and sums, unfused (this file is built with
 * -ffp-contract=off).  R keeps the rotations' rounding residue below the
 * diagonal, as in the reference.
 * fix_zero_pivots = 0 follows the source literally: a zero R(i-1,j) divides
 * by zero (NaN / Inf spread).  fix_zero_pivots = 1 is the documented
 * deviation of the GPU path: R(i-1,j) = 0 takes c = 0, s = copysign(1,
 * R(i,j)) (the rotation's limit), and both zero takes c = 1, s = 0. */
void orc_qr_givens(int m, int n, const double *A, double *Q, double *R, int fix_zero_pivots) {
    memcpy(R, A, sizeof(double) * (size_t)m * n);
    for (int c = 0; c < m; ++c) {
        memset(Q + (size_t)c * m, 0, sizeof(double) * (size_t)m);
        Q[(size_t)c * m + c] = 1.0;
    }
    for (int j = 0; j < n; ++j) {
        for (int i = m - 1; i > j; --i) {
            const double top = R[(size_t)j * m + i - 1], bot = R[(size_t)j * m + i];
            double c, s;
            if (fix_zero_pivots && top == 0.0) {
                c = bot == 0.0 ? 1.0 : 0.0;
                s = bot == 0.0 ? 0.0 : copysign(1.0, bot);
            } else {
                const double t = bot / top;
                c = 1.0 / sqrt(t * t + 1.0);
                s = t * c;
            }
            for (int k = j; k < n; ++k) {
                double *col = R + (size_t)k * m;
                const double a = col[i - 1], b = col[i];
                col[i - 1] = a * c + b * s;
                col[i] = -a * s + b * c;
            }
            double *qa = Q + (size_t)(i - 1) * m, *qb = Q + (size_t)i * m;
            for (int r = 0; r < m; ++r) {
                const double a = qa[r], b = qb[r];
                qa[r] = a * c + b * s;
                qb[r] = -a * s + b * c;
            }
        }
    }
}

/* ------------------------------------------------------------------ */
/* QR(A, Q, R, QR_GramSchmidt): QR/QR.h:36-42 with QR/__GramSchmidt.h:2-18 */
/* ------------------------------------------------------------------ */
/* Modified Gram-Schmidt on Q = A (m x n): for column j, for i < j,
 * R(i,j) = dot(Q_i, Q_j) with the four-way split dot, Q_j += Q_i * -R(i,j)
 * (__mixed, fused like the AVX path); then R(j,j) = sqrt(dot(Q_j, Q_j)) and
 * Q_j *= 1 / R(j,j) (__times).  R is m x n (ld m), zero except its upper
 * n x n triangle.  Q holds n columns (the reference orthogonalises
 * Q.cols columns of a Q presized by the caller). */
void orc_qr_gramschmidt(int m, int n, const double *A, double *Q, double *R) {
    memcpy(Q, A, sizeof(double) * (size_t)m * n);
    memset(R, 0, sizeof(double) * (size_t)m * n);
    for (int j = 0; j < n; ++j) {
        double *qj = Q + (size_t)j * m;
        for (int i = 0; i < j; ++i) {
            const double *qi = Q + (size_t)i * m;
            const double len = orc_dot(qi, qj, m);
            axpy_into(qj, qj, qi, -len, m);
            R[(size_t)j * m + i] = len;
        }
        const double len = sqrt(orc_dot(qj, qj, m));
        R[(size_t)j * m + j] = len;
        const double inv = 1.0 / len;
        for (int r = 0; r < m; ++r) qj[r] = qj[r] * inv;
    }
}

/* ------------------------------------------------------------------ */
/* LU / PLU / Solve_direct: LU/__LU.h:2-16, LU/__PLU.h:18-50,           */
/* LU/__Solve_direct.h:142-220 (the unblocked forms; the blocked ones   */
/* reorder the same updates)                                            */
/* ------------------------------------------------------------------ */
/* In place on the m x n matrix F (ld m): L (unit, strictly below the
 * diagonal) and U (on and above).  Column r first takes the updates of the
 * earlier columns (Temp[i+1:] -= L[i+1:, i] Temp[i], fused like __mixed),
 * then, with pivoting, the first row of largest |F(i, r)|, i >= r, is swapped
 * into row r across the whole row (L's columns < r and U's >= r) and
 * pos[] follows it; the column below the pivot is scaled by 1 / pivot.
 * Without pivoting a zero pivot ends the factorisation (the source's
 * `break`); with pivoting an all-zero column does.  Columns beyond the
 * square part of a wide matrix take every update (the blocked source's
 * L^{-1} A12; its unblocked loop stops after the first of them).  Returns
 * the number of columns factored. */
int orc_lu_factor(int m, int n, double *F, int pivot, int *pos) {
    const int s = imin(m, n);
    for (int i = 0; i < m; ++i) pos[i] = i;
    for (int r = 0; r < n; ++r) {
        double *col = F + (size_t)r * m;
        for (int i = 0; i < imin(r, s); ++i)
            axpy_into(col + i + 1, col + i + 1, F + (size_t)i * m + i + 1, -col[i], m - i - 1);
        /* columns right of the square part (m < n) only take the updates,
         * as the blocked source's triangular solve gives them */
        if (r >= s) continue;
        if (pivot) {
            double maxv = 0.0;
            for (int i = r; i < m; ++i) maxv = fabs(col[i]) > maxv ? fabs(col[i]) : maxv;
            if (maxv == 0.0) return r;
            int j = r;
            for (int i = r; i < m; ++i)
                if (fabs(col[i]) == maxv) { j = i; break; }
            if (j != r) {
                for (int c = 0; c < n; ++c) {
                    double *p = F + (size_t)c * m;
                    const double t = p[r]; p[r] = p[j]; p[j] = t;
                }
                const int t = pos[r]; pos[r] = pos[j]; pos[j] = t;
            }
        } else if (col[r] == 0.0) {
            return r;
        }
        const double inv = 1.0 / col[r];
        for (int i = r + 1; i < m; ++i) col[i] = col[i] * inv;
    }
    return s;
}

/* Solve_direct(A, B) (LU/__Solve_direct.h:142-220): A n x n (ld n), B and X
 * n x nrhs (ld n): PLU of a copy of A, b = B with rows permuted by pos,
 * forward substitution with the unit L, back substitution with U. */
void orc_solve_direct(int n, int nrhs, const double *A, const double *B, double *X) {
    double *F = (double *)malloc(sizeof(double) * (size_t)n * n);
    int *pos = (int *)malloc(sizeof(int) * (size_t)(n > 0 ? n : 1));
    memcpy(F, A, sizeof(double) * (size_t)n * n);
    orc_lu_factor(n, n, F, 1, pos);
    for (int j = 0; j < nrhs; ++j) {
        double *x = X + (size_t)j * n;
        const double *b = B + (size_t)j * n;
        for (int i = 0; i < n; ++i) x[i] = b[pos[i]];
        for (int k = 0; k + 1 < n; ++k)
            axpy_into(x + k + 1, x + k + 1, F + (size_t)k * n + k + 1, -x[k], n - k - 1);
        for (int r = n - 1; r > 0; --r) {
            x[r] = x[r] / F[(size_t)r * n + r];
            axpy_into(x, x, F + (size_t)r * n, -x[r], r);
        }
        if (n > 0) x[0] = x[0] / F[0];
    }
    free(F);
    free(pos);
}
