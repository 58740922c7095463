/* Test infrastructure only: a driver for the oracle's restatement of the
 * reference QR (mqr_oracle.c), built with AddressSanitizer and
 * UndefinedBehaviorSanitizer by `make sanitize`.  It runs every entry point
 * the tests use over edge-case shapes (1 x 1, wide, square, both branches of
 * QR/QR.h:7-46, a zero column) and checks ||QR - A|| / ||A||, ||Q^T Q - I||
 * and the least-squares normal equations, so that an out-of-bounds access or
 * undefined operation in the restatement stops the run.  Exit status 0 when
 * every check passes. */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "mqr_oracle.h"

static uint64_t rng_state = 88172645463325252ull;

static double uniform(void) {  /* xorshift64, in (-1, 1) */
    rng_state ^= rng_state << 13;
    rng_state ^= rng_state >> 7;
    rng_state ^= rng_state << 17;
    return (double)(rng_state >> 11) / 4503599627370496.0 - 1.0;
}

static double *alloc(size_t n) {
    double *p = calloc(n ? n : 1, sizeof(double));
    if (!p) {
        fprintf(stderr, "out of memory\n");
        exit(2);
    }
    return p;
}

/* ||Q R - A||_F / ||A||_F and ||Q^T Q - I||_F for A m x n, Q m x m, R m x n */
static void qr_errors(int m, int n, const double *A, const double *Q, const double *R,
                      double *rec, double *orth) {
    double num = 0.0, den = 0.0, o = 0.0;
    for (int j = 0; j < n; ++j)
        for (int i = 0; i < m; ++i) {
            double s = 0.0;
            for (int k = 0; k < m; ++k) s += Q[i + (int64_t)k * m] * R[k + (int64_t)j * m];
            const double d = s - A[i + (int64_t)j * m];
            num += d * d;
            den += A[i + (int64_t)j * m] * A[i + (int64_t)j * m];
        }
    for (int j = 0; j < m; ++j)
        for (int i = 0; i < m; ++i) {
            double s = 0.0;
            for (int k = 0; k < m; ++k) s += Q[k + (int64_t)i * m] * Q[k + (int64_t)j * m];
            const double d = s - (i == j ? 1.0 : 0.0);
            o += d * d;
        }
    *rec = den > 0.0 ? sqrt(num / den) : sqrt(num);
    *orth = sqrt(o);
}

static int check_qr(int m, int n, int zero_col) {
    double *A = alloc((size_t)m * n), *Q = alloc((size_t)m * m), *R = alloc((size_t)m * n);
    for (size_t e = 0; e < (size_t)m * n; ++e) A[e] = uniform();
    if (zero_col >= 0 && zero_col < n)
        for (int i = 0; i < m; ++i) A[i + (int64_t)zero_col * m] = 0.0;
    const int blocked = orc_qr_householder(m, n, A, Q, R);
    double rec, orth;
    qr_errors(m, n, A, Q, R, &rec, &orth);
    const double tol = 1e-13 * (m > n ? m : n);
    const int ok = rec < tol && orth < tol;
    printf("qr %4d x %-4d %s zero col %3d: |QR-A|/|A| %.2e |QtQ-I| %.2e %s\n", m, n,
           blocked ? "blocked  " : "unblocked", zero_col, rec, orth, ok ? "ok" : "FAIL");
    free(A);
    free(Q);
    free(R);
    return ok;
}

/* blocked factorisation + explicit Q, then least squares: A^T (b - A x) ~ 0 */
static int check_blocked_and_gels(int m, int n) {
    int cfg[6];
    orc_config(cfg);
    const int bs = cfg[5];
    const int k = m < n ? m : n;
    double *A = alloc((size_t)m * n), *VR = alloc((size_t)m * n), *Q = alloc((size_t)m * m);
    double *T = alloc((size_t)((k + bs - 1) / bs) * bs * bs + 1), *tau = alloc((size_t)k);
    double *work = alloc(orc_work_doubles(m, n));
    for (size_t e = 0; e < (size_t)m * n; ++e) A[e] = VR[e] = uniform();
    orc_geqrf_blocked(m, n, VR, T, tau, work);
    orc_orgqr_full(m, n, VR, T, Q, work);
    double *R = alloc((size_t)m * n);
    for (int j = 0; j < n; ++j)
        for (int i = 0; i <= j && i < m; ++i) R[i + (int64_t)j * m] = VR[i + (int64_t)j * m];
    double rec, orth;
    qr_errors(m, n, A, Q, R, &rec, &orth);
    const double tol = 1e-13 * (m > n ? m : n);
    int ok = rec < tol && orth < tol;
    double nres = 0.0;
    if (m >= n) {
        double *b = alloc((size_t)m), *x = alloc((size_t)n), *r = alloc((size_t)m);
        for (int i = 0; i < m; ++i) b[i] = uniform();
        orc_gels(m, n, A, b, x);
        for (int i = 0; i < m; ++i) {
            double s = b[i];
            for (int j = 0; j < n; ++j) s -= A[i + (int64_t)j * m] * x[j];
            r[i] = s;
        }
        for (int j = 0; j < n; ++j) {
            double s = 0.0;
            for (int i = 0; i < m; ++i) s += A[i + (int64_t)j * m] * r[i];
            nres = fmax(nres, fabs(s));
        }
        ok = ok && nres < 1e-11 * m;
        free(b);
        free(x);
        free(r);
    }
    printf("geqrf+orgqr+gels %4d x %-4d: |QR-A|/|A| %.2e |QtQ-I| %.2e |A^T r| %.2e %s\n", m, n,
           rec, orth, nres, ok ? "ok" : "FAIL");
    free(A);
    free(VR);
    free(Q);
    free(T);
    free(tau);
    free(work);
    free(R);
    return ok;
}

/* packed GEMM against a plain triple loop, exact on small integers */
static int check_gemm(int M, int N, int K) {
    double *A = alloc((size_t)M * K), *B = alloc((size_t)K * N), *C = alloc((size_t)M * N);
    for (size_t e = 0; e < (size_t)M * K; ++e) A[e] = (double)((int)(8 * uniform()));
    for (size_t e = 0; e < (size_t)K * N; ++e) B[e] = (double)((int)(8 * uniform()));
    orc_gemm(M, N, K, A, M, B, K, C, M);
    int ok = 1;
    for (int j = 0; j < N && ok; ++j)
        for (int i = 0; i < M; ++i) {
            double s = 0.0;
            for (int k = 0; k < K; ++k) s += A[i + (int64_t)k * M] * B[k + (int64_t)j * K];
            if (s != C[i + (int64_t)j * M]) {
                ok = 0;
                break;
            }
        }
    printf("gemm %d x %d x %d: %s\n", M, N, K, ok ? "exact" : "FAIL");
    free(A);
    free(B);
    free(C);
    return ok;
}

int main(void) {
    int ok = 1;
    const int shapes[][3] = {{1, 1, -1},    {7, 3, -1},    {3, 7, -1},   {8, 5, 0},
                             {100, 64, -1}, {97, 97, 40},  {300, 200, -1}, {250, 400, -1},
                             {260, 260, 259}};
    for (size_t s = 0; s < sizeof(shapes) / sizeof(shapes[0]); ++s)
        ok &= check_qr(shapes[s][0], shapes[s][1], shapes[s][2]);
    ok &= check_blocked_and_gels(200, 50);
    ok &= check_blocked_and_gels(257, 257);
    ok &= check_blocked_and_gels(33, 100);
    ok &= check_gemm(37, 29, 41);
    ok &= check_gemm(200, 150, 90);
    printf("%s\n", ok ? "all checks passed" : "CHECKS FAILED");
    return ok ? 0 : 1;
}
