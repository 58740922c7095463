// Drop-in check of include/mat/Mat.h: the reference's QR(A, Q, R, QR_Householder)
// call shapes (QR/QR.h:7-71, Mat.h:132-136) on MAT<double> and MAT<float>,
// in place and out of place, and LeastSquares.  Prints one line per check and
// exits non-zero on the first failure.  Inputs: mt19937_64 with a Box-Muller
// normal (bit-identical across standard libraries).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <cstdio>
#include <cstdint>
#include <random>
#include <vector>

#include "mat/Mat.h"

static int failures = 0;

static void expect(bool ok, const char *what, double value, double bound) {
    std::printf("%-48s %.3e (bound %.1e) %s\n", what, value, bound, ok ? "ok" : "FAIL");
    std::fflush(stdout);
    if (!ok) ++failures;
}

struct Normal {
    std::mt19937_64 g;
    explicit Normal(uint64_t seed) : g(seed) {}
    double operator()() {
        const double u1 = ((g() >> 11) + 0.5) * (1.0 / 9007199254740992.0);
        const double u2 = ((g() >> 11) + 0.5) * (1.0 / 9007199254740992.0);
        return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
    }
};

template <typename T>
MAT<T> randn(int m, int n, uint64_t seed) {
    Normal rng(seed);
    MAT<T> A(m, n);
    for (int64_t i = 0; i < (int64_t)m * n; ++i) A.data[i] = (T)rng();
    return A;
}

// ||Q R - A||_F / ||A||_F, ||Q^T Q - I||_F, max |R(r, c)| below the diagonal
template <typename T>
void qr_errors(const MAT<T> &A, const MAT<T> &Q, const MAT<T> &R, double *rec, double *orth,
               double *low) {
    const int m = A.rows, n = A.cols;
    double num = 0, den = 0, o = 0, l = 0;
    for (int c = 0; c < n; ++c)
        for (int r = 0; r < m; ++r) {
            double s = 0;
            for (int k = 0; k < m; ++k) s += (double)Q(r, k) * (double)R(k, c);
            const double a = (double)A(r, c);
            num += (s - a) * (s - a);
            den += a * a;
            if (r > c) l = std::max(l, std::fabs((double)R(r, c)));
        }
    for (int c = 0; c < m; ++c)
        for (int r = 0; r < m; ++r) {
            double s = 0;
            for (int k = 0; k < m; ++k) s += (double)Q(k, r) * (double)Q(k, c);
            s -= (r == c) ? 1.0 : 0.0;
            o += s * s;
        }
    *rec = std::sqrt(num / den);
    *orth = std::sqrt(o);
    *low = l;
}

int main() {
    const double u = 1.1102230246251565e-16;
    {
        // blocked-size FP64 call, the reference's dispatch condition
        // min(m, n) > 2 * BS on both ISAs
        MAT<double> A = randn<double>(400, 260, 1), Q, R;
        QR_init(A, Q, R);
        QR(A, Q, R, QR_Householder);
        double rec, orth, low;
        qr_errors(A, Q, R, &rec, &orth, &low);
        expect(Q.rows == 400 && Q.cols == 400 && R.rows == 400 && R.cols == 260,
               "QR(A,Q,R,QR_Householder) shapes", 0.0, 0.0);
        expect(rec < 64 * 400 * u, "FP64 400x260 ||QR-A||/||A||", rec, 64 * 400 * u);
        expect(orth < 64 * 400 * u, "FP64 400x260 ||Q^TQ-I||", orth, 64 * 400 * u);
        expect(low == 0.0, "FP64 400x260 R zero below the diagonal", low, 0.0);
    }
    {
        // the template form, a small (unblocked-size) matrix, wide shape
        MAT<double> A = randn<double>(40, 70, 2), Q, R;
        QR<QR_Householder>(A, Q, R);
        double rec, orth, low;
        qr_errors(A, Q, R, &rec, &orth, &low);
        expect(rec < 64 * 70 * u, "QR<QR_Householder> 40x70 ||QR-A||/||A||", rec, 64 * 70 * u);
        expect(orth < 64 * 70 * u, "QR<QR_Householder> 40x70 ||Q^TQ-I||", orth, 64 * 70 * u);
    }
    {
        // in place: A is R (QR/QR.h:12-13)
        MAT<double> A = randn<double>(300, 300, 3), A0(A), Q;
        QR(A, Q, A, QR_Householder);
        double rec, orth, low;
        qr_errors(A0, Q, A, &rec, &orth, &low);
        expect(rec < 64 * 300 * u, "in place 300x300 ||QR-A||/||A||", rec, 64 * 300 * u);
    }
    {
        // FP32 matrices (computed in FP64, converted like the A -> R copy)
        MAT<float> A = randn<float>(200, 150, 4), Q, R;
        QR_init(A, Q, R);
        QR(A, Q, R, QR_Householder);
        double rec, orth, low;
        qr_errors(A, Q, R, &rec, &orth, &low);
        const double uf = 5.960464477539063e-08;
        expect(rec < 64 * uf, "FP32 200x150 ||QR-A||/||A||", rec, 64 * uf);
        expect(orth < 64 * uf * 14, "FP32 200x150 ||Q^TQ-I||", orth, 64 * uf * 14);
    }
    {
        // least squares, plain and mixed: the normal equations hold
        MAT<double> A = randn<double>(900, 300, 5), b = randn<double>(900, 1, 6), x, x32;
        LeastSquares(A, b, x);
        mqr_ir_info_t info = LeastSquares(A, b, x32, MQR_PREC_FP32);
        double na = 0, g = 0, g32 = 0, nx = 0, dx = 0;
        for (int64_t i = 0; i < (int64_t)A.rows * A.cols; ++i) na += A.data[i] * A.data[i];
        // r = b - A x and A^T r
        std::vector<double> r(A.rows), r32(A.rows);
        for (int i = 0; i < A.rows; ++i) {
            double s = b(i, 0), s32 = b(i, 0);
            for (int c = 0; c < A.cols; ++c) {
                s -= A(i, c) * x(c, 0);
                s32 -= A(i, c) * x32(c, 0);
            }
            r[i] = s;
            r32[i] = s32;
        }
        for (int c = 0; c < A.cols; ++c) {
            double s = 0, s32 = 0;
            for (int i = 0; i < A.rows; ++i) {
                s += A(i, c) * r[i];
                s32 += A(i, c) * r32[i];
            }
            g += s * s;
            g32 += s32 * s32;
            nx += x(c, 0) * x(c, 0);
            dx += (x(c, 0) - x32(c, 0)) * (x(c, 0) - x32(c, 0));
        }
        double nb = 0;
        for (int i = 0; i < A.rows; ++i) nb += b(i, 0) * b(i, 0);
        const double eta = std::sqrt(g) / (na * std::sqrt(nx) + std::sqrt(na * nb));
        const double eta32 = std::sqrt(g32) / (na * std::sqrt(nx) + std::sqrt(na * nb));
        expect(eta < 1e-15, "LeastSquares FP64 ||A^T r|| / (||A||(||A|| ||x|| + ||b||))", eta, 1e-15);
        expect(eta32 < 1e-15, "LeastSquares FP32+IR same", eta32, 1e-15);
        expect(info.converged == 1, "LeastSquares FP32+IR converged", info.iterations, 30);
        expect(std::sqrt(dx / nx) < 1e-12, "LeastSquares FP32+IR vs FP64 x", std::sqrt(dx / nx), 1e-12);
    }
    {
        // the other QR_type values on the same switch (QR/QR.h:26-42)
        MAT<double> A = randn<double>(60, 35, 7), Q, R;
        QR_init(A, Q, R);
        QR(A, Q, R, QR_Givens);
        double rec, orth, low;
        qr_errors(A, Q, R, &rec, &orth, &low);
        expect(Q.rows == 60 && Q.cols == 60 && R.rows == 60 && R.cols == 35, "QR_Givens shapes",
               0.0, 0.0);
        expect(rec < 64 * 60 * u, "QR_Givens 60x35 ||QR-A||/||A||", rec, 64 * 60 * u);
        expect(orth < 64 * 60 * u, "QR_Givens 60x35 ||Q^TQ-I||", orth, 64 * 60 * u);
        MAT<double> G, S;
        QR<QR_GramSchmidt>(A, G, S);
        double num = 0, den = 0, o = 0, dmin = 1e300, lowgs = 0;
        for (int c = 0; c < 35; ++c) {
            dmin = std::min(dmin, S(c, c));
            for (int r = 0; r < 60; ++r) {
                double acc = 0;
                for (int k = 0; k < 35; ++k) acc += G(r, k) * S(k, c);
                num += (acc - A(r, c)) * (acc - A(r, c));
                den += A(r, c) * A(r, c);
                if (r > c) lowgs = std::max(lowgs, std::fabs(S(r, c)));
            }
            for (int k = 0; k < 35; ++k) {
                double acc = 0;
                for (int r = 0; r < 60; ++r) acc += G(r, k) * G(r, c);
                const double e = acc - (k == c ? 1.0 : 0.0);
                o += e * e;
            }
        }
        expect(G.rows == 60 && G.cols == 35 && S.rows == 60 && S.cols == 35,
               "QR_GramSchmidt shapes (Q m x n)", 0.0, 0.0);
        expect(std::sqrt(num / den) < 64 * 60 * u, "QR_GramSchmidt ||QR-A||/||A||",
               std::sqrt(num / den), 64 * 60 * u);
        expect(std::sqrt(o) < 64 * 60 * u, "QR_GramSchmidt ||Q^TQ-I||", std::sqrt(o), 64 * 60 * u);
        expect(dmin > 0 && lowgs == 0.0, "QR_GramSchmidt R upper, positive diagonal", dmin, 0.0);
    }
    {
        // operator* on the GPU: the reference's own reconstruction check
        // Q * R - A (Mul/Mul.h:205-235), exact on integer-valued operands
        MAT<double> A = randn<double>(300, 170, 8), Q, R;
        QR(A, Q, R, QR_Householder);
        MAT<double> QR_ = Q * R;
        MAT<double> D = QR_ - A;
        double num = 0, den = 0;
        for (int64_t i = 0; i < (int64_t)A.rows * A.cols; ++i) {
            num += D.data[i] * D.data[i];
            den += A.data[i] * A.data[i];
        }
        expect(QR_.rows == 300 && QR_.cols == 170, "operator* shape", 0.0, 0.0);
        expect(std::sqrt(num / den) < 64 * 300 * u, "||Q*R - A|| / ||A|| via operator*",
               std::sqrt(num / den), 64 * 300 * u);
        MAT<double> I1(37, 53), I2(53, 29);
        MAT<float> F1(37, 53), F2(53, 29);
        std::mt19937_64 g(9);
        for (int i = 0; i < 37 * 53; ++i) F1.data[i] = (float)(I1.data[i] = (double)(int)(g() % 17) - 8);
        for (int i = 0; i < 53 * 29; ++i) F2.data[i] = (float)(I2.data[i] = (double)(int)(g() % 17) - 8);
        MAT<double> P = I1 * I2;
        MAT<float> PF = F1 * F2;
        MAT<double> PM = I1 * F2;  // mixed element types: FP64, result double
        double worst = 0;
        for (int c = 0; c < 29; ++c)
            for (int r = 0; r < 37; ++r) {
                double s = 0;
                for (int k = 0; k < 53; ++k) s += I1(r, k) * I2(k, c);
                worst = std::max({worst, std::fabs(P(r, c) - s), std::fabs((double)PF(r, c) - s),
                                  std::fabs(PM(r, c) - s)});
            }
        expect(worst == 0.0, "operator* exact on integers (FP64, FP32, mixed)", worst, 0.0);
    }
    {
        // data model: _T, dagger, diagonal, swap_rows / swap_cols, __BlockMat,
        // scalar arithmetic (Base/__Matrix.h, Base/__Block.h)
        MAT<double> A = randn<double>(5, 3, 10);
        MAT<double> T = _T(A), H = A.dagger(), T2 = A.transpose();
        bool ok = T.rows == 3 && T.cols == 5;
        for (int r = 0; r < 5; ++r)
            for (int c = 0; c < 3; ++c) ok = ok && T(c, r) == A(r, c) && H(c, r) == A(r, c) && T2(c, r) == A(r, c);
        expect(ok, "_T / dagger / transpose", 0.0, 0.0);
        double d[4] = {1, 2, 3, 4};
        MAT<double> Dg;
        Dg.diagonal(d, 4);
        ok = Dg.rows == 4 && Dg.cols == 4;
        for (int r = 0; r < 4; ++r)
            for (int c = 0; c < 4; ++c) ok = ok && Dg(r, c) == (r == c ? d[r] : 0.0);
        MAT<double> S(A);
        S.swap_rows(0, 4);
        S.swap_cols(0, 2);
        for (int c = 0; c < 3; ++c)
            for (int r = 0; r < 5; ++r) {
                const int rr = r == 0 ? 4 : r == 4 ? 0 : r, cc = c == 0 ? 2 : c == 2 ? 0 : c;
                ok = ok && S(r, c) == A(rr, cc);
            }
        expect(ok, "diagonal / swap_rows / swap_cols", 0.0, 0.0);
        MAT<double> B = randn<double>(6, 6, 11), B0(B);
        __BlockMat<double> V(B, 1, 2, 4, 3);
        ok = V(0, 0) == B(1, 2) && V[2][3] == B(4, 4) && V._rows == 6;
        V.swap_rows(0, 3, 0, 2);
        for (int c = 0; c < 6; ++c)
            for (int r = 0; r < 6; ++r) {
                double want = B0(r, c);
                if ((c == 2 || c == 3) && r == 1) want = B0(4, c);
                if ((c == 2 || c == 3) && r == 4) want = B0(1, c);
                ok = ok && B(r, c) == want;
            }
        MAT<double> E = 2.0 * A + A / 2.0 - A * 0.5;
        for (int i = 0; i < 15; ++i) ok = ok && E.data[i] == 2.0 * A.data[i] + A.data[i] / 2.0 - A.data[i] * 0.5;
        expect(ok, "__BlockMat view / swap_rows, scalar * / + -", 0.0, 0.0);
    }
    {
        // binary MAT files (Mat-5.2/_extra/__File.h): int32 rows, int32 cols,
        // then the elements; round trips through every load / save form
        const std::string f = "/tmp/mat_api_test.mat", fr = "/tmp/mat_api_test.raw";
        MAT<double> A = randn<double>(7, 4, 12), B, C;
        save(f, A);
        bool ok = size(f) == 8 + 8 * 28;
        load(f, B);
        ok = ok && B.rows == 7 && B.cols == 4 && std::memcmp(A.data, B.data, 8 * 28) == 0;
        save(fr, A.data, 28);
        load(fr, C, 7, 4);
        ok = ok && size(fr) == 8 * 28 && std::memcmp(A.data, C.data, 8 * 28) == 0;
        float fx[3] = {1.5f, -2.f, 3.25f}, fy[3] = {0, 0, 0};
        save(fr, fx, 3);
        load(fr, fy, 3);
        ok = ok && fy[0] == 1.5f && fy[1] == -2.f && fy[2] == 3.25f;
        bool threw = false;
        try {
            MAT<double> Z;
            load(std::string("/tmp/does_not_exist_mat_api.mat"), Z);
        } catch (const mat::Error &e) {
            threw = e.status == MQR_ERR_IO;
        }
        std::remove(f.c_str());
        std::remove(fr.c_str());
        expect(ok && threw, "MAT file save / load / size round trips", 0.0, 0.0);
    }
    {
        // LU, PLU, PLU_getP, LU_getLU and Solve_direct (LU/)
        MAT<double> A = randn<double>(150, 150, 13), L, U, P, Lc, Uc;
        std::vector<int> pos(150);
        LU_init(A, L, U);
        const int info = PLU(A, pos.data(), L, U);
        P.init(150, 150);
        PLU_getP(pos.data(), P);
        MAT<double> PLU_ = P * (L * U);
        double num = 0, den = 0, up = 0;
        for (int c = 0; c < 150; ++c)
            for (int r = 0; r < 150; ++r) {
                num += (PLU_(r, c) - A(r, c)) * (PLU_(r, c) - A(r, c));
                den += A(r, c) * A(r, c);
                if (r > c) up = std::max(up, std::fabs(U(r, c)));
                if (r < c) up = std::max(up, std::fabs(L(r, c)));
            }
        expect(info == 0 && up == 0.0, "PLU: L unit lower, U upper", up, 0.0);
        expect(std::sqrt(num / den) < 64 * 150 * u, "PLU: ||P L U - A|| / ||A||",
               std::sqrt(num / den), 64 * 150 * u);
        MAT<double> D = randn<double>(80, 80, 14);
        for (int i = 0; i < 80; ++i) D(i, i) += 30.0;
        LU(D, L, U);
        MAT<double> F(80, 80);
        for (int c = 0; c < 80; ++c)
            for (int r = 0; r < 80; ++r) F(r, c) = r > c ? L(r, c) : U(r, c);
        Lc.init(80, 80);
        Uc.init(80, 80);
        LU_getLU(F, Lc, Uc);
        MAT<double> LUd = Lc * Uc;
        num = den = 0;
        for (int i = 0; i < 80 * 80; ++i) {
            num += (LUd.data[i] - D.data[i]) * (LUd.data[i] - D.data[i]);
            den += D.data[i] * D.data[i];
        }
        expect(std::sqrt(num / den) < 64 * 80 * u, "LU (no pivoting) + LU_getLU: ||L U - A||",
               std::sqrt(num / den), 64 * 80 * u);
        MAT<double> B = randn<double>(150, 3, 15);
        MAT<double> X = Solve_direct(A, B);
        MAT<double> AX = A * X;
        num = den = 0;
        for (int i = 0; i < 150 * 3; ++i) {
            num += (AX.data[i] - B.data[i]) * (AX.data[i] - B.data[i]);
            den += B.data[i] * B.data[i];
        }
        expect(std::sqrt(num / den) < 1e-10, "Solve_direct: ||A X - B|| / ||B||",
               std::sqrt(num / den), 1e-10);
    }
    std::printf("%s\n", failures ? "FAILED" : "ALL PASSED");
    return failures ? 1 : 0;
}
