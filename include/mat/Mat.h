// Mat.h -- header-only C++ API of the MI355X QR library, kept call-compatible
// with the reference's dense-factorisation surface:
//
//   MAT<T>                      Base/__Matrix.h:2-174 (column-major owning matrix)
//   enum QR_type                Mat.h:43-47
//   QR_init(A, Q, R)            Mat.h:132-136
//   QR(A, Q, R, QR_Householder) QR/QR.h:7-46
//   QR<QR_Householder>(A, Q, R) QR/QR.h:48-71
//   LeastSquares(A, b, x)       no reference counterpart (new on the same path)
//
// Every Householder QR call runs on the GPU through the C ABI of
// include/mat_qr.h (libmatqr.so): A is copied to device memory, factored in
// FP64 (blocked compact-WY Householder, the reference's reflector
// conventions), Q (m x m) and R (m x n, zero below the diagonal) are copied
// back.  There is no CPU fallback and no runtime backend choice: a failure of
// the device library is reported as a mat::Error exception from these host
// helpers (the C ABI itself never throws).  QR_Givens and QR_GramSchmidt are
// not part of this library and raise mat::Error.
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

    MAT transpose() const {
        MAT t(cols, rows);
        for (int c = 0; c < cols; ++c)
            for (int r = 0; r < rows; ++r) t.data[(size_t)t.rows * r + c] = data[(size_t)rows * c + r];
        return t;
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

}  // namespace detail
}  // namespace mat

// QR/QR.h:7-46.  Q is (re)sized to m x m and R to m x n.
template <typename t1, typename t2>
inline void QR(MAT<t1> &A, MAT<t2> &Q, MAT<t2> &R, QR_type type) {
    switch (type) {
        case QR_Householder:
            mat::detail::qr_householder(A, Q, R);
            break;
        default:
            throw mat::Error(MQR_ERR_UNSUPPORTED, "QR: only QR_Householder runs on the GPU");
    }
}

// QR/QR.h:48-71
template <QR_type type, typename t1, typename t2>
inline void QR(MAT<t1> &A, MAT<t2> &Q, MAT<t2> &R) {
    static_assert(type == QR_Householder, "only QR_Householder runs on the GPU");
    mat::detail::qr_householder(A, Q, R);
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
