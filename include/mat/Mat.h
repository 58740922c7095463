// Mat.h -- header-only C++ API of the MI355X QR library, kept call-compatible
// with the reference's dense-factorisation surface:
//
//   MAT<T>                      Base/__Matrix.h:2-174 (column-major owning matrix)
//   enum QR_type                Mat.h:43-47
//   QR_init(A, Q, R)            Mat.h:132-136
//   QR(A, Q, R, QR_Householder) QR/QR.h:7-46
//   QR<QR_Householder>(A, Q, R) QR/QR.h:48-71
//   LeastSquares(A, b, x)       no reference counterpart (new on the same path)
//   operator*(MAT, MAT)         Mul/Mul.h:205-235, on the GPU (MFMA GEMM)
//   _T, _H, scalar * / + -      Base/__Matrix.h:177-259 (host)
//   __BlockMat<T>               Base/__Block.h:2-47 (host view)
//   load / save / size          Mat-5.2/_extra/__File.h:1-66 (binary MAT files)
//
// Every Householder QR call runs on the GPU through the C ABI of
// include/mat_qr.h (libmatqr.so): A is copied to device memory, factored in
// FP64 (blocked compact-WY Householder, the reference's reflector
// conventions), Q (m x m) and R (m x n, zero below the diagonal) are copied
// back.  There is no CPU fallback and no runtime backend choice: a failure of
// the device library is reported as a mat::Error exception from these host
// helpers (the C ABI itself never throws).  QR_Givens and QR_GramSchmidt run
// on the GPU as well (Q m x m for Givens, m x n for Gram-Schmidt).
//
// Element type: the factorisation computes in FP64; MAT<float> inputs and
// outputs are converted on the host, like the reference's type-converting
// A -> R copy (QR/QR.h:11-13).
//
// One handle per thread and device is created on first use and kept until
// the thread exits (mat::handle()).
#ifndef MAT_GPU_MAT_H
#define MAT_GPU_MAT_H

#include <algorithm>
#include <cstring>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <utility>
#include <vector>

#include "../mat_qr.h"

namespace mat {

class Error : public std::runtime_error {
public:
    int status;
    Error(int s, const std::string &where)
        : std::runtime_error(where + ": " + mqr_strerror(s) + " (status " + std::to_string(s) + ")"),
          status(s) {}
};

inline void check(int s, const char *where) {
    if (s != MQR_SUCCESS) throw Error(s, where);
}

// Per-thread library handle on device 0 (MAT_QR_DEVICE_INDEX selects another).
#ifndef MAT_QR_DEVICE_INDEX
#define MAT_QR_DEVICE_INDEX 0
#endif
struct HandleOwner {
    mqr_handle_t h = nullptr;
    HandleOwner() { check(mqr_create(&h, MAT_QR_DEVICE_INDEX), "mqr_create"); }
    ~HandleOwner() {
        if (h) mqr_destroy(h);
    }
    HandleOwner(const HandleOwner &) = delete;
    HandleOwner &operator=(const HandleOwner &) = delete;
};
inline mqr_handle_t handle() {
    thread_local HandleOwner owner;
    return owner.h;
}

}  // namespace mat

// ---------------------------------------------------------------------------
// MAT<T>: column-major owning matrix, element (r, c) at data[rows * c + r]
// (Base/__Matrix.h:144-146).  Same members and semantics as the reference:
// init() reallocates only on a shape change, operator= takes B's buffer
// (B is passed by value), operator& exchanges two matrices.
// ---------------------------------------------------------------------------
template <typename T>
class MAT {
public:
    int rows;
    int cols;
    T *data;

    MAT() : rows(0), cols(0), data(nullptr) {}
    MAT(int r, int c) : rows(r), cols(c), data(new T[(size_t)r * c]) {}
    template <class t>
    MAT(int r, int c, t *p) : rows(r), cols(c), data(new T[(size_t)r * c]) {
        const size_t s = (size_t)r * c;
        for (size_t i = 0; i < s; ++i) data[i] = (T)p[i];
    }
    MAT(MAT &a) : rows(a.rows), cols(a.cols), data(new T[(size_t)a.rows * a.cols]) {
        std::memcpy(data, a.data, sizeof(T) * (size_t)rows * cols);
    }
    MAT(MAT &&a) : rows(a.rows), cols(a.cols), data(a.data) { a.data = nullptr; }
    template <class t>
    MAT(const MAT<t> &a) : rows(a.rows), cols(a.cols), data(new T[(size_t)a.rows * a.cols]) {
        const size_t s = (size_t)rows * cols;
        for (size_t i = 0; i < s; ++i) data[i] = (T)a.data[i];
    }
    ~MAT() {
        delete[] data;
        data = nullptr;
    }

    void init(int r, int c) {
        if (rows != r || cols != c) {
            rows = r;
            cols = c;
            delete[] data;
            data = new T[(size_t)r * c];
        }
    }
    template <typename t>
    void init(int r, int c, t *p) {
        init(r, c);
        const size_t s = (size_t)r * c;
        for (size_t i = 0; i < s; ++i) data[i] = (T)p[i];
    }
    void iden(int a) {
        init(a, a);
        std::fill(data, data + (size_t)a * a, T(0));
        for (int i = 0; i < a; ++i) data[(size_t)i * cols + i] = T(1);
    }
    void fill(T a) { std::fill(data, data + (size_t)rows * cols, a); }

    // zero matrix with a on the diagonal (Base/__Matrix.h:97-103): the
    // reference writes rows entries; min(rows, cols) are written here, so
    // a non-square matrix is not written past its end
    void diagonal(T *a) {
        std::fill(data, data + (size_t)rows * cols, T(0));
        for (int i = 0; i < std::min(rows, cols); ++i) data[(size_t)i * rows + i] = a[i];
    }
    // size x size, a on the diagonal, converted to T (:104-112)
    template <typename t>
    void diagonal(t *a, int size) {
        init(size, size);
        std::fill(data, data + (size_t)size * size, T(0));
        for (int i = 0; i < size; ++i) data[(size_t)i * rows + i] = (T)a[i];
    }
    // exchange two rows / two columns in place (:117-135)
    void swap_rows(int r1, int r2) {
        for (int c = 0; c < cols; ++c) std::swap(data[(size_t)c * rows + r1], data[(size_t)c * rows + r2]);
    }
    void swap_cols(int c1, int c2) {
        std::swap_ranges(data + (size_t)rows * c1, data + (size_t)rows * (c1 + 1),
                         data + (size_t)rows * c2);
    }

    T *operator[](int c) const { return data + (size_t)rows * c; }
    T &operator()(int r, int c) const { return data[(size_t)rows * c + r]; }

    void operator=(MAT<T> B) {
        rows = B.rows;
        cols = B.cols;
        std::swap(data, B.data);
    }
    void operator&(MAT &B) {
        std::swap(rows, B.rows);
        std::swap(cols, B.cols);
        std::swap(data, B.data);
    }

    MAT transpose() const;
    // conjugate transpose; for the real types here the transpose (:203-206)
    MAT dagger() const;
};

// _T(A): a new cols x rows matrix holding A^T (Base/__Matrix.h:177-186)
template <typename t>
MAT<t> _T(const MAT<t> &A) {
    MAT<t> n(A.cols, A.rows);
    for (int c = 0; c < A.cols; ++c)
        for (int r = 0; r < A.rows; ++r) n.data[(size_t)n.rows * r + c] = A.data[(size_t)A.rows * c + r];
    return n;
}
// _H(A): the conjugate transpose, the transpose for real element types (:191-200)
template <typename t>
MAT<t> _H(const MAT<t> &A) {
    return _T(A);
}
template <typename T>
MAT<T> MAT<T>::transpose() const {
    return _T(*this);
}
template <typename T>
MAT<T> MAT<T>::dagger() const {
    return _H(*this);
}

// Elementwise scalar and matrix arithmetic (Base/__Matrix.h:209-259).  Host
// loops: these are O(mn) data-model operations next to the GPU products.
template <class t>
MAT<t> operator*(t a, const MAT<t> &A) {
    MAT<t> n(A.rows, A.cols);
    for (size_t i = 0; i < (size_t)A.rows * A.cols; ++i) n.data[i] = A.data[i] * a;
    return n;
}
template <class t>
MAT<t> operator*(const MAT<t> &A, t a) {
    return a * A;
}
template <class t>
MAT<t> operator/(const MAT<t> &A, t a) {
    return (t(1) / a) * A;
}
template <typename t1, typename t2>
auto operator+(const MAT<t1> &A, const MAT<t2> &B) {
    MAT<decltype(std::declval<t1>() + std::declval<t2>())> n(A.rows, A.cols);
    for (size_t i = 0; i < (size_t)A.rows * A.cols; ++i) n.data[i] = A.data[i] + B.data[i];
    return n;
}
template <typename t1, typename t2>
auto operator-(const MAT<t1> &A, const MAT<t2> &B) {
    MAT<decltype(std::declval<t1>() - std::declval<t2>())> n(A.rows, A.cols);
    for (size_t i = 0; i < (size_t)A.rows * A.cols; ++i) n.data[i] = A.data[i] - B.data[i];
    return n;
}

// __BlockMat<T>: non-owning r x c view of a MAT at (m, n), leading dimension
// A.rows (Base/__Block.h:2-47).
template <typename T>
class __BlockMat {
public:
    int rows;
    int cols;
    int _rows;
    T *data;
    __BlockMat() : rows(0), cols(0), _rows(0), data(nullptr) {}
    __BlockMat(const MAT<T> &A, int m, int n, int r, int c)
        : rows(r), cols(c), _rows(A.rows), data(A.data + (size_t)A.rows * n + m) {}
    void set(const MAT<T> &A, int m, int n, int r, int c) {
        rows = r;
        cols = c;
        _rows = A.rows;
        data = A.data + (size_t)A.rows * n + m;
    }
    T *operator[](int c) { return data + (size_t)_rows * c; }
    T &operator()(int r, int c) { return data[(size_t)_rows * c + r]; }
    // exchange rows r1 and r2 over the columns [c1, c2) (:36-46)
    void swap_rows(int r1, int r2, int c1, int c2) {
        for (int c = c1; c < c2; ++c) std::swap(data[(size_t)c * _rows + r1], data[(size_t)c * _rows + r2]);
    }
};

// Mat.h:43-47
enum QR_type { QR_Householder, QR_Givens, QR_GramSchmidt };

// Mat.h:132-136
template <typename t1, typename t2>
void QR_init(MAT<t1> &A, MAT<t2> &Q, MAT<t2> &R) {
    Q.init(A.rows, A.rows);
    R.init(A.rows, A.cols);
}

namespace mat {
namespace detail {

template <typename T>
const double *as_f64(const MAT<T> &M, std::vector<double> &buf) {
    if constexpr (std::is_same<T, double>::value) {
        (void)buf;
        return M.data;
    } else {
        buf.assign(M.data, M.data + (size_t)M.rows * M.cols);
        return buf.data();
    }
}

template <typename T>
void from_f64(MAT<T> &M, const std::vector<double> &src) {
    const size_t s = (size_t)M.rows * M.cols;
    for (size_t i = 0; i < s; ++i) M.data[i] = (T)src[i];
}

// Householder QR on the GPU: Q m x m, R m x n (zero below the diagonal).
template <typename t1, typename t2>
void qr_householder(const MAT<t1> &A, MAT<t2> &Q, MAT<t2> &R) {
    const int m = A.rows, n = A.cols;
    std::vector<double> abuf;
    const double *a = as_f64(A, abuf);
    // A may alias R (in place, QR/QR.h:12-13): the C call reads A before
    // writing R, and Q/R are written through separate host buffers.
    std::vector<double> q((size_t)m * m), r((size_t)m * n);
    check(mqr_qr_householder_host(handle(), m, n, a, q.data(), r.data()),
          "mqr_qr_householder_host");
    Q.init(m, m);
    R.init(m, n);
    if constexpr (std::is_same<t2, double>::value) {
        std::memcpy(Q.data, q.data(), sizeof(double) * q.size());
        std::memcpy(R.data, r.data(), sizeof(double) * r.size());
    } else {
        from_f64(Q, q);
        from_f64(R, r);
    }
}

// QR_Givens (Q m x m, R m x n) and QR_GramSchmidt (Q m x n, R m x n) on the
// GPU through mqr_qr_host; A may alias R or Q (it is read before either is
// written).
template <typename t1, typename t2>
void qr_other(const MAT<t1> &A, MAT<t2> &Q, MAT<t2> &R, mqr_qr_type_t type) {
    const int m = A.rows, n = A.cols;
    const int qc = type == MQR_QR_GIVENS ? m : n;
    std::vector<double> abuf;
    const double *a = as_f64(A, abuf);
    std::vector<double> q((size_t)m * qc), r((size_t)m * n);
    check(mqr_qr_host(handle(), type, m, n, a, q.data(), r.data()), "mqr_qr_host");
    Q.init(m, qc);
    R.init(m, n);
    from_f64(Q, q);
    from_f64(R, r);
}

}  // namespace detail
}  // namespace mat

// Mul/Mul.h:205-235: C = A * B (A.rows x B.cols) on the GPU's MFMA GEMM
// (mqr_dgemm_host / mqr_sgemm_host).  float x float multiplies in FP32;
// every other element type through FP64 and converted to the result type.
template <typename t1, typename t2>
auto operator*(const MAT<t1> &A, const MAT<t2> &B) {
    using t3 = decltype(std::declval<t1>() * std::declval<t2>());
    MAT<t3> C(A.rows, B.cols);
    if (A.cols != B.rows) throw mat::Error(-3, "operator*: A.cols != B.rows");
    if constexpr (std::is_same<t1, float>::value && std::is_same<t2, float>::value) {
        mat::check(mqr_sgemm_host(mat::handle(), A.rows, B.cols, A.cols, A.data, B.data, C.data),
                   "mqr_sgemm_host");
    } else if constexpr (std::is_same<t1, double>::value && std::is_same<t2, double>::value) {
        mat::check(mqr_dgemm_host(mat::handle(), A.rows, B.cols, A.cols, A.data, B.data, C.data),
                   "mqr_dgemm_host");
    } else {
        std::vector<double> a, b, c((size_t)A.rows * B.cols);
        const double *pa = mat::detail::as_f64(A, a), *pb = mat::detail::as_f64(B, b);
        mat::check(mqr_dgemm_host(mat::handle(), A.rows, B.cols, A.cols, pa, pb, c.data()),
                   "mqr_dgemm_host");
        mat::detail::from_f64(C, c);
    }
    return C;
}

// ---------------------------------------------------------------------------
// Binary MAT files (Mat-5.2/_extra/__File.h:1-66): int32 rows, int32 cols,
// then the elements in storage order (column-major here), no type tag.
// Failures raise mat::Error(MQR_ERR_IO) where the reference reads garbage.
// ---------------------------------------------------------------------------
// file size in bytes (:3-10)
inline long size(const std::string &name) {
    int64_t b = -1;
    mat::check(mqr_file_size(name.c_str(), &b), "size");
    return (long)b;
}
// len elements without a header (:12-19)
template <class T>
void load(const std::string &name, T *data, int len) {
    mat::check(mqr_raw_load(name.c_str(), data, sizeof(T) * (size_t)len), "load");
}
// header and elements; A takes the file's shape (:20-34)
template <class T>
void load(const std::string &name, MAT<T> &A) {
    int32_t r = 0, c = 0;
    mat::check(mqr_mat_load_header(name.c_str(), &r, &c), "load");
    A.init(r, c);
    mat::check(mqr_mat_load(name.c_str(), r, c, A.data, sizeof(T)), "load");
}
// r x c elements without a header (:35-44)
template <class T>
void load(const std::string &name, MAT<T> &A, int r, int c) {
    A.init(r, c);
    mat::check(mqr_raw_load(name.c_str(), A.data, sizeof(T) * (size_t)r * c), "load");
}
// len elements without a header (:48-55)
template <class T>
void save(const std::string &name, T *data, int len) {
    mat::check(mqr_raw_save(name.c_str(), data, sizeof(T) * (size_t)len), "save");
}
// header and elements (:56-66)
template <class T>
void save(const std::string &name, MAT<T> &A) {
    mat::check(mqr_mat_save(name.c_str(), A.rows, A.cols, A.data, sizeof(T)), "save");
}

// QR/QR.h:7-46.  Q is (re)sized to m x m and R to m x n.
template <typename t1, typename t2>
inline void QR(MAT<t1> &A, MAT<t2> &Q, MAT<t2> &R, QR_type type) {
    switch (type) {
        case QR_Householder:
            mat::detail::qr_householder(A, Q, R);
            break;
        case QR_Givens:
            mat::detail::qr_other(A, Q, R, MQR_QR_GIVENS);
            break;
        case QR_GramSchmidt:
            mat::detail::qr_other(A, Q, R, MQR_QR_GRAMSCHMIDT);
            break;
        default:
            throw mat::Error(MQR_ERR_UNSUPPORTED, "QR: unknown QR_type");
    }
}

// QR/QR.h:48-71
template <QR_type type, typename t1, typename t2>
inline void QR(MAT<t1> &A, MAT<t2> &Q, MAT<t2> &R) {
    QR(A, Q, R, type);
}

// ---------------------------------------------------------------------------
// LU, PLU and the direct solve on the GPU (LU/__LU.h, LU/__PLU.h,
// LU/__Solve_direct.h; LU_init, Mat.h:128-131).  L is m x m (unit lower: the
// multipliers in its first min(m, n) columns, the identity beyond), U is
// m x n (zero below the diagonal), so A = L U, or row pos[i] of A = row i of
// L U for PLU.  Where the reference stops at an exactly zero pivot, the GPU
// factorisation continues LAPACK-style; the 1-based column of the first such
// pivot is returned (0: none).
// ---------------------------------------------------------------------------
template <typename t1, typename t2>
void LU_init(MAT<t1> &A, MAT<t2> &L, MAT<t2> &U) {
    L.init(A.rows, A.rows);
    U.init(A.rows, A.cols);
}

namespace mat {
namespace detail {
template <typename t1, typename t2>
int lu_split(const MAT<t1> &A, int pivot, int *pos, MAT<t2> &L, MAT<t2> &U) {
    const int m = A.rows, n = A.cols, s = std::min(m, n);
    std::vector<double> abuf, f((size_t)m * n);
    const double *a = as_f64(A, abuf);
    std::vector<int32_t> p((size_t)std::max(m, 1));
    int32_t info = 0;
    check(mqr_lu_host(handle(), pivot, m, n, a, f.data(), p.data(), &info), "mqr_lu_host");
    L.init(m, m);
    U.init(m, n);
    for (int c = 0; c < m; ++c)
        for (int r = 0; r < m; ++r)
            L(r, c) = (t2)(r == c ? 1.0 : (c < s && r > c) ? f[(size_t)c * m + r] : 0.0);
    for (int c = 0; c < n; ++c)
        for (int r = 0; r < m; ++r) U(r, c) = (t2)(r <= c ? f[(size_t)c * m + r] : 0.0);
    if (pos)
        for (int i = 0; i < m; ++i) pos[i] = p[(size_t)i];
    return info;
}
}  // namespace detail
}  // namespace mat

// LU/__LU.h:86-103, no pivoting
template <typename t1, typename t2>
int LU(MAT<t1> &A, MAT<t2> &L, MAT<t2> &U) {
    return mat::detail::lu_split(A, 0, nullptr, L, U);
}

// LU/__PLU.h:153-187, partial pivoting: pos (A.rows entries) receives the
// row permutation
template <typename t1, typename t2>
int PLU(MAT<t1> &A, int *pos, MAT<t2> &L, MAT<t2> &U) {
    return mat::detail::lu_split(A, 1, pos, L, U);
}

// LU/__PLU.h:189-195: P with P(pos[i], i) = 1, so that A = P L U
template <typename t>
void PLU_getP(int *pos, MAT<t> &P) {
    P.fill(t(0));
    for (int i = 0; i < P.rows; ++i) P(pos[i], i) = t(1);
}

// LU/__LU.h:105-130: L and U from a combined L \ U matrix
template <typename t1, typename t2>
void LU_getLU(MAT<t1> &F, MAT<t2> &L, MAT<t2> &U) {
    const int s = std::min(F.rows, F.cols);
    for (int c = 0; c < L.cols; ++c)
        for (int r = 0; r < L.rows; ++r)
            L(r, c) = (t2)(r == c ? 1 : (c < s && r > c) ? F(r, c) : 0);
    for (int c = 0; c < U.cols; ++c)
        for (int r = 0; r < U.rows; ++r)
            U(r, c) = (t2)((r <= c || c >= s) && r < F.rows && c < F.cols ? F(r, c) : 0);
}

// LU/__Solve_direct.h:175-220: X = A^{-1} B by PLU on the GPU
template <typename T>
MAT<T> Solve_direct(MAT<T> A, MAT<T> &B) {
    MAT<T> X(B.rows, B.cols);
    if (A.rows != A.cols || A.rows != B.rows)
        throw mat::Error(-1, "Solve_direct: A must be square with A.rows == B.rows");
    std::vector<double> abuf, bbuf, x((size_t)B.rows * B.cols);
    const double *a = mat::detail::as_f64(A, abuf), *b = mat::detail::as_f64(B, bbuf);
    mat::check(mqr_solve_direct_host(mat::handle(), A.rows, B.cols, a, b, x.data()),
               "mqr_solve_direct_host");
    mat::detail::from_f64(X, x);
    return X;
}

// ---------------------------------------------------------------------------
// Least squares x = argmin ||A x - b|| (m >= n), no reference counterpart.
// prec: MQR_PREC_FP64 (plain FP64 QR solve) or MQR_PREC_FP32 (FP32 trailing
// updates + FP64 refinement to the FP64 solution).  x is resized to n x 1.
// Returns the refinement record (zeroed for the plain solve).
// ---------------------------------------------------------------------------
inline mqr_ir_info_t LeastSquares(MAT<double> &A, MAT<double> &b, MAT<double> &x,
                                  mqr_precision_t prec = MQR_PREC_FP64) {
    if (b.rows != A.rows || b.cols != 1)
        throw mat::Error(-3, "LeastSquares: b must be an m x 1 matrix");
    mqr_ir_info_t info;
    std::memset(&info, 0, sizeof(info));
    x.init(A.cols, 1);
    mat::check(mqr_gels_host(mat::handle(), prec, A.rows, A.cols, A.data, b.data, x.data,
                             prec == MQR_PREC_FP64 ? nullptr : &info),
               "mqr_gels_host");
    return info;
}

#endif  // MAT_GPU_MAT_H
