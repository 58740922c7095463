// C ABI entry points (include/mat_qr.h): handle, workspace, argument checks,
// and the host-memory conveniences the header-only MAT<T> API calls.
#include <algorithm>
#include <cstring>
#include <new>
#include <type_traits>
#include <vector>

#include "mqr_internal.h"

namespace mqr {

int geqrf_f64(mqr_handle_t hd, int64_t m, int64_t n, double *A, int64_t lda, double *tau,
              double *T, int64_t ldt);
int ormqr_f64(mqr_handle_t hd, bool trans, int64_t m, int64_t nrhs, int64_t k, const double *VR,
              int64_t ldvr, const double *T, int64_t ldt, double *C, int64_t ldc);
int orgqr_f64(mqr_handle_t hd, int64_t m, int64_t nq, int64_t k, const double *VR, int64_t ldvr,
              const double *T, int64_t ldt, double *Q, int64_t ldq);
int trsm_upper_f64(mqr_handle_t hd, bool trans, int64_t n, int64_t nrhs, const double *R,
                   int64_t ldr, double *B, int64_t ldb);
int geqrf_batched(mqr_handle_t hd, bool f64, int64_t batch, const int32_t *m, const int32_t *n,
                  const int64_t *lda, void *const *A, void *const *tau, int tile);
int batched_plan_create(mqr_handle_t hd, bool f64, int64_t batch, const int32_t *m,
                        const int32_t *n, const int64_t *lda, void *const *A, void *const *tau,
                        int tile, mqr_batched_plan_st **out, int per_cu, int order);
int batched_plan_run(mqr_handle_t hd, mqr_batched_plan_st *pl);
void batched_plan_destroy(mqr_batched_plan_st *pl);
int batched_plan_device(const mqr_batched_plan_st *pl);
int qr_backward(mqr_handle_t hd, bool f64, int64_t batch, int64_t m, int64_t n, const void *Q,
                int64_t ldq, int64_t sq, const void *R, int64_t ldr, int64_t sr, const void *dQ,
                int64_t lddq, int64_t sdq, const void *dR, int64_t lddr, int64_t sdr, void *dA,
                int64_t ldda, int64_t sda);
size_t givens_ws_bytes(int64_t m, int64_t n);
bool givens_supported(int64_t m, int64_t n);
int qr_givens_f64(mqr_handle_t hd, int64_t m, int64_t n, const double *A, int64_t lda, double *Q,
                  int64_t ldq, double *R, int64_t ldr);
int qr_gramschmidt_f64(mqr_handle_t hd, int64_t m, int64_t n, const double *A, int64_t lda,
                       double *Q, int64_t ldq, double *R, int64_t ldr);
int getrf_f64(mqr_handle_t hd, int64_t m, int64_t n, double *A, int64_t lda, int32_t *ipiv,
              int32_t *info, int32_t *scratch_piv, int32_t *scratch_info);
int getrs_f64(mqr_handle_t hd, int64_t n, int64_t nrhs, const double *LU, int64_t ldlu,
              const int32_t *ipiv, double *B, int64_t ldb);
int batched_noalloc_init(int device);
int geqrf_batched_noalloc(mqr_handle_t hd, bool f64, int64_t batch, const int32_t *m,
                          const int32_t *n, const int64_t *lda, void *const *A, void *const *tau,
                          int tile);
int gels_mixed(mqr_handle_t hd, mqr_precision_t prec, int64_t m, int64_t n, const double *A,
               int64_t lda, const double *b, double *x, const mqr_ir_opts_t *opts,
               mqr_ir_info_t *info);

namespace {
thread_local char g_last_hip[256] = {0};

// Makes the handle's device current for the duration of a call.
struct DeviceGuard {
    int prev = -1;
    bool ok = true;
    explicit DeviceGuard(int dev) {
        if (hipGetDevice(&prev) != hipSuccess) prev = -1;
        if (prev != dev) ok = hipSetDevice(dev) == hipSuccess;
    }
    ~DeviceGuard() {
        if (prev >= 0) (void)hipSetDevice(prev);
    }
};
}  // namespace

void set_last_hip_error(hipError_t e, const char *where) {
    std::snprintf(g_last_hip, sizeof(g_last_hip), "%s: %s", where, hipGetErrorString(e));
}

int ws_reserve(mqr_handle_t h, Workspace &ws, size_t bytes) {
    if (ws.bytes >= bytes) return MQR_SUCCESS;
    MQR_HIP(hipStreamSynchronize(h->stream));
    MQR_HIP(hipStreamSynchronize(h->side_stream));
    MQR_HIP(hipStreamSynchronize(h->far_stream));
    if (ws.ptr) {
        MQR_HIP(hipFree(ws.ptr));
        ws.ptr = nullptr;
        ws.bytes = 0;
    }
    const size_t want = ((bytes + bytes / 4) + (2u << 20) - 1) & ~(size_t)((2u << 20) - 1);
    if (hipMalloc(&ws.ptr, want) != hipSuccess) {
        ws.ptr = nullptr;
        (void)hipGetLastError();
        return MQR_ERR_ALLOC;
    }
    ws.bytes = want;
    return MQR_SUCCESS;
}

}  // namespace mqr

using namespace mqr;

#define MQR_GUARD(h)                                  \
    if (!(h)) return -1;                              \
    DeviceGuard guard_((h)->device);                  \
    if (!guard_.ok) return MQR_ERR_NO_DEVICE;         \
    ::mqr::t_prof = (h)->prof.on ? &(h)->prof : nullptr

// Orders a call's work (queued on the handle's own stream) after what the
// caller queued on its stream before the call, and the caller's later work
// after the call's: the caller sees ordinary stream-ordered semantics while
// the kernels run on a non-blocking stream (dependent launches on the legacy
// null stream cost about twice as much).
struct StreamBridge {
    mqr_handle_t h;
    bool active;
    int status = MQR_SUCCESS;
    explicit StreamBridge(mqr_handle_t hd) : h(hd), active(hd->user_stream != hd->stream) {
        if (!active) return;
        if (hipEventRecord(h->ev_in, h->user_stream) != hipSuccess ||
            hipStreamWaitEvent(h->stream, h->ev_in, 0) != hipSuccess) {
            ::mqr::set_last_hip_error(hipGetLastError(), "stream bridge (in)");
            status = MQR_ERR_HIP;
        }
    }
    ~StreamBridge() {
        if (!active) return;
        (void)hipEventRecord(h->ev_out, h->stream);
        (void)hipStreamWaitEvent(h->user_stream, h->ev_out, 0);
    }
};

#define MQR_BRIDGE(h)              \
    StreamBridge bridge_(h);       \
    if (bridge_.status) return bridge_.status

extern "C" {

const char *mqr_strerror(int status) {
    if (status == MQR_SUCCESS) return "success";
    if (status < 0) return "invalid argument (the status is minus its 1-based position)";
    switch (status) {
        case MQR_ERR_NO_DEVICE: return "no HIP device available (or device index out of range)";
        case MQR_ERR_HIP: return g_last_hip[0] ? g_last_hip : "HIP runtime error";
        case MQR_ERR_ALLOC: return "device memory allocation failed";
        case MQR_ERR_NOT_CONVERGED: return "iterative refinement did not converge";
        case MQR_ERR_UNSUPPORTED: return "request not supported by this build";
        case MQR_ERR_NOT_INITIALIZED: return "handle not initialised";
        case MQR_ERR_INTERNAL: return "internal error";
        case MQR_ERR_IO:
            return "file could not be opened, read or written, or its size or header does not "
                   "match the request";
        case MQR_ERR_TOO_LARGE:
            return "request exceeds the fixed device scratch of the allocation-free entry point "
                   "(at most MQR_NOALLOC_MAX_BATCH = 32768 matrices)";
        default: return "unknown status";
    }
}

int mqr_get_device_count(int *count) {
    if (!count) return -1;
    int c = 0;
    if (hipGetDeviceCount(&c) != hipSuccess) {
        (void)hipGetLastError();
        *count = 0;
        return MQR_ERR_NO_DEVICE;
    }
    *count = c;
    return c > 0 ? MQR_SUCCESS : MQR_ERR_NO_DEVICE;
}

int mqr_create(mqr_handle_t *handle, int device) {
    if (!handle) return -1;
    *handle = nullptr;
    int c = 0;
    if (hipGetDeviceCount(&c) != hipSuccess || c <= 0) {
        (void)hipGetLastError();
        return MQR_ERR_NO_DEVICE;
    }
    if (device < 0 || device >= c) return -2;
    DeviceGuard guard(device);
    if (!guard.ok) return MQR_ERR_NO_DEVICE;
    mqr_handle_st *h = new (std::nothrow) mqr_handle_st();
    if (!h) return MQR_ERR_ALLOC;
    h->device = device;
    if (hipStreamCreateWithFlags(&h->own_stream, hipStreamNonBlocking) != hipSuccess) {
        delete h;
        return MQR_ERR_HIP;
    }
    h->stream = h->own_stream;
    // second stream and events of the look-ahead schedule (default priority)
    h->user_stream = h->own_stream;
    if (hipStreamCreateWithFlags(&h->side_stream, hipStreamNonBlocking) != hipSuccess ||
        hipStreamCreateWithFlags(&h->far_stream, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&h->ev_fork, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&h->ev_join, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&h->ev_side, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&h->ev_in, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&h->ev_out, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&h->ev_host, hipEventDisableTiming) != hipSuccess) {
        (void)hipGetLastError();
        for (hipEvent_t e : {h->ev_fork, h->ev_join, h->ev_side, h->ev_in, h->ev_out, h->ev_host})
            if (e) (void)hipEventDestroy(e);
        if (h->side_stream) (void)hipStreamDestroy(h->side_stream);
        if (h->far_stream) (void)hipStreamDestroy(h->far_stream);
        (void)hipStreamDestroy(h->own_stream);
        delete h;
        return MQR_ERR_HIP;
    }
    int cus = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) == hipSuccess &&
        cus > 0)
        h->cu_count = cus;
    if (hipHostMalloc((void **)&h->host_pinned, 64 * sizeof(double), hipHostMallocDefault) !=
        hipSuccess) {
        for (hipEvent_t e : {h->ev_fork, h->ev_join, h->ev_side, h->ev_in, h->ev_out, h->ev_host}) (void)hipEventDestroy(e);
        (void)hipStreamDestroy(h->side_stream);
        (void)hipStreamDestroy(h->far_stream);
        (void)hipStreamDestroy(h->own_stream);
        delete h;
        return MQR_ERR_ALLOC;
    }
    // this device's address of the allocation-free entry point's scratch
    if (batched_noalloc_init(device) != MQR_SUCCESS) {
        (void)hipGetLastError();
        (void)hipHostFree(h->host_pinned);
        for (hipEvent_t e : {h->ev_fork, h->ev_join, h->ev_side, h->ev_in, h->ev_out, h->ev_host}) (void)hipEventDestroy(e);
        (void)hipStreamDestroy(h->side_stream);
        (void)hipStreamDestroy(h->far_stream);
        (void)hipStreamDestroy(h->own_stream);
        delete h;
        return MQR_ERR_HIP;
    }
    *handle = h;
    return MQR_SUCCESS;
}

int mqr_destroy(mqr_handle_t h) {
    MQR_GUARD(h);
    (void)hipStreamSynchronize(h->stream);
    (void)hipStreamSynchronize(h->side_stream);
    (void)hipStreamSynchronize(h->far_stream);
    if (h->user_stream != h->stream) (void)hipStreamSynchronize(h->user_stream);
    if (h->ws.ptr) (void)hipFree(h->ws.ptr);
    if (h->ws_ir.ptr) (void)hipFree(h->ws_ir.ptr);
    if (h->ws_vec.ptr) (void)hipFree(h->ws_vec.ptr);
    if (h->ws_lu.ptr) (void)hipFree(h->ws_lu.ptr);
    if (h->host_pinned) (void)hipHostFree(h->host_pinned);
    for (hipEvent_t e : {h->ev_fork, h->ev_join, h->ev_side, h->ev_in, h->ev_out, h->ev_host})
        if (e) (void)hipEventDestroy(e);
    if (h->side_stream) (void)hipStreamDestroy(h->side_stream);
    if (h->far_stream) (void)hipStreamDestroy(h->far_stream);
    if (h->own_stream) (void)hipStreamDestroy(h->own_stream);
    delete h;
    return MQR_SUCCESS;
}

int mqr_set_stream(mqr_handle_t h, void *stream) {
    if (!h) return -1;
    h->user_stream = (hipStream_t)stream;
    return MQR_SUCCESS;
}

int mqr_get_stream(mqr_handle_t h, void **stream) {
    if (!h) return -1;
    if (!stream) return -2;
    *stream = (void *)h->user_stream;
    return MQR_SUCCESS;
}

int mqr_set_block_size(mqr_handle_t h, int nb) {
    if (!h) return -1;
    // 16..64 in steps of 16: panels of nb; 128, 192, 256: outer blocks of
    // 64-column panels (two-level blocking, T blocks nb wide)
    if (nb < 16 || nb > kMaxOuterNb || nb % 16 != 0 || (nb > kMaxNb && nb % kMaxNb != 0)) return -2;
    h->nb = nb;
    return MQR_SUCCESS;
}

int mqr_get_block_size(mqr_handle_t h, int *nb) {
    if (!h) return -1;
    if (!nb) return -2;
    *nb = h->nb;
    return MQR_SUCCESS;
}

int mqr_synchronize(mqr_handle_t h) {
    MQR_GUARD(h);
    MQR_HIP(hipStreamSynchronize(h->stream));
    MQR_HIP(hipStreamSynchronize(h->side_stream));
    MQR_HIP(hipStreamSynchronize(h->far_stream));
    if (h->user_stream != h->stream) MQR_HIP(hipStreamSynchronize(h->user_stream));
    return MQR_SUCCESS;
}

int mqr_malloc(mqr_handle_t h, void **ptr, size_t bytes) {
    MQR_GUARD(h);
    if (!ptr) return -2;
    *ptr = nullptr;
    if (bytes == 0) return MQR_SUCCESS;
    if (hipMalloc(ptr, bytes) != hipSuccess) {
        (void)hipGetLastError();
        *ptr = nullptr;
        return MQR_ERR_ALLOC;
    }
    return MQR_SUCCESS;
}

int mqr_free(mqr_handle_t h, void *ptr) {
    MQR_GUARD(h);
    if (ptr) MQR_HIP(hipFree(ptr));
    return MQR_SUCCESS;
}

int mqr_memcpy_htod(mqr_handle_t h, void *dst, const void *src, size_t bytes) {
    MQR_GUARD(h);
    if (bytes && !dst) return -2;
    if (bytes && !src) return -3;
    if (bytes) {
        MQR_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, h->user_stream));
        MQR_HIP(hipStreamSynchronize(h->user_stream));
    }
    return MQR_SUCCESS;
}

int mqr_memcpy_dtoh(mqr_handle_t h, void *dst, const void *src, size_t bytes) {
    MQR_GUARD(h);
    if (bytes && !dst) return -2;
    if (bytes && !src) return -3;
    if (bytes) {
        MQR_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, h->user_stream));
        MQR_HIP(hipStreamSynchronize(h->user_stream));
    }
    return MQR_SUCCESS;
}

int mqr_dgemm(mqr_handle_t h, char transa, char transb, int64_t m, int64_t n, int64_t k,
              double alpha, const double *A, int64_t lda, const double *B, int64_t ldb,
              double beta, double *C, int64_t ldc) {
    mqr::Range rg_("mqr_dgemm");
    if (!h) return -1;
    const bool ta = transa == 'T' || transa == 't', tb = transb == 'T' || transb == 't';
    if (!ta && transa != 'N' && transa != 'n') return -2;
    if (!tb && transb != 'N' && transb != 'n') return -3;
    if (m < 0 || m > INT32_MAX) return -4;
    if (n < 0 || n > INT32_MAX) return -5;
    if (k < 0 || k > INT32_MAX) return -6;
    if (m * k > 0 && !A) return -8;
    if (lda < std::max<int64_t>(1, ta ? k : m)) return -9;
    if (k * n > 0 && !B) return -10;
    if (ldb < std::max<int64_t>(1, tb ? n : k)) return -11;
    if (m * n > 0 && !C) return -13;
    if (ldc < std::max<int64_t>(1, m)) return -14;
    if (m == 0 || n == 0) return MQR_SUCCESS;
    MQR_GUARD(h);
    MQR_BRIDGE(h);
    if (k == 0 || alpha == 0.0) return scale_f64(m, n, beta, C, ldc, h->stream);
    GemmF64 g;
    g.M = (int)m; g.N = (int)n; g.K = (int)k;
    g.ta = ta; g.tb = tb;
    g.A = A; g.lda = lda; g.B = B; g.ldb = ldb;
    g.C = C; g.ldc = ldc; g.alpha = alpha; g.beta = beta;
    return gemm_f64(g, h->stream);
}

static int sgemm_common(mqr_handle_t h, int terms, char transa, char transb, int64_t m,
                        int64_t n, int64_t k, float alpha, const float *A, int64_t lda,
                        const float *B, int64_t ldb, float beta, float *C, int64_t ldc) {
    if (!h) return -1;
    const bool ta = transa == 'T' || transa == 't', tb = transb == 'T' || transb == 't';
    if (!ta && transa != 'N' && transa != 'n') return -2;
    if (!tb && transb != 'N' && transb != 'n') return -3;
    if (m < 0 || m > INT32_MAX) return -4;
    if (n < 0 || n > INT32_MAX) return -5;
    if (k < 0 || k > INT32_MAX) return -6;
    if (m * k > 0 && !A) return -8;
    if (lda < std::max<int64_t>(1, ta ? k : m)) return -9;
    if (k * n > 0 && !B) return -10;
    if (ldb < std::max<int64_t>(1, tb ? n : k)) return -11;
    if (m * n > 0 && !C) return -13;
    if (ldc < std::max<int64_t>(1, m)) return -14;
    if (m == 0 || n == 0) return MQR_SUCCESS;
    if (k == 0) return MQR_ERR_UNSUPPORTED;
    MQR_GUARD(h);
    MQR_BRIDGE(h);
    GemmF32 g;
    g.M = (int)m; g.N = (int)n; g.K = (int)k;
    g.ta = ta; g.tb = tb;
    g.A = A; g.lda = lda; g.B = B; g.ldb = ldb;
    g.C = C; g.ldc = ldc; g.alpha = alpha; g.beta = beta;
    g.bf16_terms = terms;
    return gemm_f32(g, h->stream);
}

int mqr_sgemm(mqr_handle_t h, char transa, char transb, int64_t m, int64_t n, int64_t k,
              float alpha, const float *A, int64_t lda, const float *B, int64_t ldb, float beta,
              float *C, int64_t ldc) {
    mqr::Range rg_("mqr_sgemm");
    return sgemm_common(h, 0, transa, transb, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc);
}

int mqr_sgemm_bf16(mqr_handle_t h, int terms, char transa, char transb, int64_t m, int64_t n,
                   int64_t k, float alpha, const float *A, int64_t lda, const float *B,
                   int64_t ldb, float beta, float *C, int64_t ldc) {
    mqr::Range rg_("mqr_sgemm_bf16");
    if (!h) return -1;
    if (terms != 1 && terms != 3) return -2;
    const int st = sgemm_common(h, terms, transa, transb, m, n, k, alpha, A, lda, B, ldb, beta, C,
                                ldc);
    return st < 0 ? st - 1 : st;  // argument positions shift by one
}

} // extern "C"

// C = A B on host buffers: device copies of A, B and C, one MFMA GEMM, C back.
template <typename T>
static int gemm_host(mqr_handle_t h, int64_t m, int64_t n, int64_t k, const T *A, const T *B,
                     T *C) {
    if (!h) return -1;
    if (m < 0 || m > INT32_MAX) return -2;
    if (n < 0 || n > INT32_MAX) return -3;
    if (k < 0 || k > INT32_MAX) return -4;
    if (m * k > 0 && !A) return -5;
    if (k * n > 0 && !B) return -6;
    if (m * n > 0 && !C) return -7;
    if (m == 0 || n == 0) return MQR_SUCCESS;
    MQR_GUARD(h);
    MQR_BRIDGE(h);
    T *d = nullptr;
    const size_t na = (size_t)(m * k), nbb = (size_t)(k * n), nc = (size_t)(m * n);
    if (hipMalloc(&d, sizeof(T) * (na + nbb + nc)) != hipSuccess) {
        (void)hipGetLastError();
        return MQR_ERR_ALLOC;
    }
    T *dA = d, *dB = d + na, *dC = d + na + nbb;
    int st = MQR_SUCCESS;
    if ((na && hipMemcpyAsync(dA, A, sizeof(T) * na, hipMemcpyHostToDevice, h->stream) != hipSuccess) ||
        (nbb && hipMemcpyAsync(dB, B, sizeof(T) * nbb, hipMemcpyHostToDevice, h->stream) != hipSuccess))
        st = MQR_ERR_HIP;
    if (st == MQR_SUCCESS) {
        if (k == 0) {
            st = hipMemsetAsync(dC, 0, sizeof(T) * nc, h->stream) == hipSuccess ? MQR_SUCCESS
                                                                             : MQR_ERR_HIP;
        } else if constexpr (std::is_same<T, double>::value) {
            GemmF64 g;
            g.M = (int)m; g.N = (int)n; g.K = (int)k;
            g.A = dA; g.lda = m; g.B = dB; g.ldb = k; g.C = dC; g.ldc = m;
            g.alpha = 1.0; g.beta = 0.0;
            st = gemm_f64(g, h->stream);
        } else {
            GemmF32 g;
            g.M = (int)m; g.N = (int)n; g.K = (int)k;
            g.A = dA; g.lda = m; g.B = dB; g.ldb = k; g.C = dC; g.ldc = m;
            g.alpha = 1.f; g.beta = 0.f;
            st = gemm_f32(g, h->stream);
        }
    }
    if (st == MQR_SUCCESS &&
        (hipMemcpyAsync(C, dC, sizeof(T) * nc, hipMemcpyDeviceToHost, h->stream) != hipSuccess ||
         hipStreamSynchronize(h->stream) != hipSuccess))
        st = MQR_ERR_HIP;
    (void)hipStreamSynchronize(h->stream);
    (void)hipFree(d);
    return st;
}

// Records the block size T is written with; a later mqr_dorgqr / mqr_dormqr
// reading the same T with another block size is rejected.
static void remember_t_layout(mqr_handle_t h, const double *T) {
    for (auto &e : h->t_layout)
        if (e.T == T) {
            e.nb = h->nb;
            return;
        }
    h->t_layout[h->t_layout_next] = {T, h->nb};
    h->t_layout_next = (h->t_layout_next + 1) % (int)(sizeof(h->t_layout) / sizeof(h->t_layout[0]));
}

static bool t_layout_mismatch(mqr_handle_t h, const double *T) {
    for (const auto &e : h->t_layout)
        if (e.T == T) return e.nb != h->nb;
    return false;
}

extern "C" {

int mqr_dgemm_host(mqr_handle_t h, int64_t m, int64_t n, int64_t k, const double *A,
                   const double *B, double *C) {
    mqr::Range rg_("mqr_dgemm_host");
    return gemm_host<double>(h, m, n, k, A, B, C);
}

int mqr_sgemm_host(mqr_handle_t h, int64_t m, int64_t n, int64_t k, const float *A,
                   const float *B, float *C) {
    mqr::Range rg_("mqr_sgemm_host");
    return gemm_host<float>(h, m, n, k, A, B, C);
}

int mqr_dgeqrf(mqr_handle_t h, int64_t m, int64_t n, double *A, int64_t lda, double *tau,
               double *T, int64_t ldt) {
    mqr::Range rg_("mqr_dgeqrf");
    if (!h) return -1;
    if (m < 0) return -2;
    if (n < 0) return -3;
    const int64_t k = std::min(m, n);
    if (m * n > 0 && !A) return -4;
    if (lda < std::max<int64_t>(1, m)) return -5;
    if (k > 0 && !tau) return -6;
    if (k > 0 && !T) return -7;
    if (ldt < std::min<int64_t>(h->nb, std::max<int64_t>(k, 1))) return -8;
    if (k == 0) return MQR_SUCCESS;
    if (m > INT32_MAX || n > INT32_MAX) return MQR_ERR_UNSUPPORTED;
    MQR_GUARD(h);
    MQR_BRIDGE(h);
    remember_t_layout(h, T);
    return geqrf_f64(h, m, n, A, lda, tau, T, ldt);
}

int mqr_dorgqr(mqr_handle_t h, int64_t m, int64_t nq, int64_t k, const double *VR, int64_t ldvr,
               const double *T, int64_t ldt, double *Q, int64_t ldq) {
    mqr::Range rg_("mqr_dorgqr");
    if (!h) return -1;
    if (m < 0) return -2;
    if (nq < 0 || nq > m) return -3;
    if (k < 0 || k > m) return -4;
    if (k > 0 && !VR) return -5;
    if (ldvr < std::max<int64_t>(1, m)) return -6;
    if (k > 0 && !T) return -7;
    if (ldt < std::min<int64_t>(h->nb, std::max<int64_t>(k, 1))) return -8;
    if (k > 0 && t_layout_mismatch(h, T)) return -7;
    if (m * nq > 0 && !Q) return -9;
    if (ldq < std::max<int64_t>(1, m)) return -10;
    if (m * nq == 0) return MQR_SUCCESS;
    MQR_GUARD(h);
    MQR_BRIDGE(h);
    return orgqr_f64(h, m, nq, k, VR, ldvr, T, ldt, Q, ldq);
}

int mqr_dormqr(mqr_handle_t h, char trans, int64_t m, int64_t nrhs, int64_t k, const double *VR,
               int64_t ldvr, const double *T, int64_t ldt, double *C, int64_t ldc) {
    mqr::Range rg_("mqr_dormqr");
    if (!h) return -1;
    const bool tr = trans == 'T' || trans == 't';
    if (!tr && trans != 'N' && trans != 'n') return -2;
    if (m < 0) return -3;
    if (nrhs < 0) return -4;
    if (k < 0 || k > m) return -5;
    if (k > 0 && !VR) return -6;
    if (ldvr < std::max<int64_t>(1, m)) return -7;
    if (k > 0 && !T) return -8;
    if (ldt < std::min<int64_t>(h->nb, std::max<int64_t>(k, 1))) return -9;
    if (k > 0 && t_layout_mismatch(h, T)) return -8;
    if (m * nrhs > 0 && !C) return -10;
    if (ldc < std::max<int64_t>(1, m)) return -11;
    if (m * nrhs == 0 || k == 0) return MQR_SUCCESS;
    MQR_GUARD(h);
    MQR_BRIDGE(h);
    return ormqr_f64(h, tr, m, nrhs, k, VR, ldvr, T, ldt, C, ldc);
}

int mqr_dtrsm_upper(mqr_handle_t h, char trans, int64_t n, int64_t nrhs, const double *R,
                    int64_t ldr, double *B, int64_t ldb) {
    mqr::Range rg_("mqr_dtrsm_upper");
    if (!h) return -1;
    const bool tr = trans == 'T' || trans == 't';
    if (!tr && trans != 'N' && trans != 'n') return -2;
    if (n < 0) return -3;
    if (nrhs < 0) return -4;
    if (n > 0 && !R) return -5;
    if (ldr < std::max<int64_t>(1, n)) return -6;
    if (n * nrhs > 0 && !B) return -7;
    if (ldb < std::max<int64_t>(1, n)) return -8;
    if (n * nrhs == 0) return MQR_SUCCESS;
    MQR_GUARD(h);
    MQR_BRIDGE(h);
    return trsm_upper_f64(h, tr, n, nrhs, R, ldr, B, ldb);
}

int mqr_dgels(mqr_handle_t h, int64_t m, int64_t n, int64_t nrhs, double *A, int64_t lda,
              double *B, int64_t ldb) {
    mqr::Range rg_("mqr_dgels");
    if (!h) return -1;
    if (m < 0) return -2;
    if (n < 0 || n > m) return -3;
    if (nrhs < 0) return -4;
    if (m * n > 0 && !A) return -5;
    if (lda < std::max<int64_t>(1, m)) return -6;
    if (m * nrhs > 0 && !B) return -7;
    if (ldb < std::max<int64_t>(1, m)) return -8;
    if (n == 0 || nrhs == 0) return MQR_SUCCESS;
    MQR_GUARD(h);
    MQR_BRIDGE(h);
    const int nb = h->nb;
    double *tauT = nullptr;
    // tau (n) and T (nb x n) live in the refinement workspace
    MQR_CHECK(ws_reserve(h, h->ws_ir, sizeof(double) * ((size_t)n * (nb + 1) + 64)));
    tauT = (double *)h->ws_ir.ptr;
    double *T = tauT + n;
    MQR_CHECK(geqrf_rhs_f64(h, m, n, A, lda, tauT, T, nb, B, ldb, nrhs));
    return trsm_upper_f64(h, false, n, nrhs, A, lda, B, ldb);
}

int mqr_gels_mixed(mqr_handle_t h, mqr_precision_t prec, int64_t m, int64_t n, const double *A,
                   int64_t lda, const double *b, double *x, const mqr_ir_opts_t *opts,
                   mqr_ir_info_t *info) {
    mqr::Range rg_("mqr_gels_mixed");
    if (!h) return -1;
    if (prec < MQR_PREC_FP64 || prec > MQR_PREC_BF16) return -2;
    if (m < 0) return -3;
    if (n < 0 || n > m) return -4;
    if (m * n > 0 && !A) return -5;
    if (lda < std::max<int64_t>(1, m)) return -6;
    if (m > 0 && !b) return -7;
    if (n > 0 && !x) return -8;
    if (info) std::memset(info, 0, sizeof(*info));
    if (n == 0) return MQR_SUCCESS;
    MQR_GUARD(h);
    MQR_BRIDGE(h);
    return gels_mixed(h, prec, m, n, A, lda, b, x, opts, info);
}

// argument checks shared by the batched entry points (LAPACK-style -i)
static int check_batched(mqr_handle_t h, mqr_dtype_t dtype, int64_t batch, const int32_t *m,
                         const int32_t *n, const int64_t *lda, void *const *A, void *const *tau,
                         int tile) {
    if (!h) return -1;
    if (dtype != MQR_F64 && dtype != MQR_F32) return -2;
    if (batch < 0 || batch > INT32_MAX) return -3;
    if (batch > 0 && (!m || !n || !lda || !A || !tau)) return !m ? -4 : !n ? -5 : !lda ? -6 : !A ? -7 : -8;
    for (int64_t i = 0; i < batch; ++i) {
        if (m[i] < 0 || m[i] > 512) return -4;
        if (n[i] < 0) return -5;
        if (lda[i] < std::max<int64_t>(1, m[i])) return -6;
        if ((int64_t)m[i] * n[i] > 0 && !A[i]) return -7;
        if (std::min(m[i], n[i]) > 0 && !tau[i]) return -8;
    }
    if (tile < -1 || tile > 1) return -9;
    return MQR_SUCCESS;
}

int mqr_geqrf_batched_uneven(mqr_handle_t h, mqr_dtype_t dtype, int64_t batch, const int32_t *m,
                             const int32_t *n, const int64_t *lda, void *const *A,
                             void *const *tau, int tile) {
    mqr::Range rg_("mqr_geqrf_batched_uneven");
    const int a = check_batched(h, dtype, batch, m, n, lda, A, tau, tile);
    if (a) return a;
    if (batch == 0) return MQR_SUCCESS;
    MQR_GUARD(h);
    MQR_BRIDGE(h);
    return geqrf_batched(h, dtype == MQR_F64, batch, m, n, lda, A, tau, tile);
}

int mqr_geqrf_batched_uneven_noalloc(mqr_handle_t h, mqr_dtype_t dtype, int64_t batch,
                                     const int32_t *m, const int32_t *n, const int64_t *lda,
                                     void *const *A, void *const *tau, int tile) {
    mqr::Range rg_("mqr_geqrf_batched_uneven_noalloc");
    if (!h) return -1;
    if (batch > MQR_NOALLOC_MAX_BATCH) return MQR_ERR_TOO_LARGE;
    const int a = check_batched(h, dtype, batch, m, n, lda, A, tau, tile);
    if (a) return a;
    if (batch == 0) return MQR_SUCCESS;
    MQR_GUARD(h);
    MQR_BRIDGE(h);
    return geqrf_batched_noalloc(h, dtype == MQR_F64, batch, m, n, lda, A, tau, tile);
}

int mqr_batched_plan_create(mqr_handle_t h, mqr_dtype_t dtype, int64_t batch, const int32_t *m,
                            const int32_t *n, const int64_t *lda, void *const *A,
                            void *const *tau, int tile, mqr_batched_plan_t *plan) {
    mqr::Range rg_("mqr_batched_plan_create");
    const int a = check_batched(h, dtype, batch, m, n, lda, A, tau, tile);
    if (a) return a;
    if (!plan) return -10;
    *plan = nullptr;
    MQR_GUARD(h);
    return batched_plan_create(h, dtype == MQR_F64, batch, m, n, lda, A, tau, tile, plan, -1, -1);
}

int mqr_batched_plan_create_sched(mqr_handle_t h, mqr_dtype_t dtype, int64_t batch,
                                  const int32_t *m, const int32_t *n, const int64_t *lda,
                                  void *const *A, void *const *tau,
                                  const mqr_batched_schedule_t *sched, mqr_batched_plan_t *plan) {
    mqr::Range rg_("mqr_batched_plan_create_sched");
    if (!sched) return -9;
    if (sched->nb != 0 && sched->nb != 8 && sched->nb != 16) return -9;
    if (sched->per_cu < -1 || sched->order < -1 || sched->order > 2) return -9;
    const int tile = sched->nb == 8 ? 0 : sched->nb == 16 ? 1 : -1;
    const int a = check_batched(h, dtype, batch, m, n, lda, A, tau, tile);
    if (a) return a;
    if (!plan) return -10;
    *plan = nullptr;
    MQR_GUARD(h);
    return batched_plan_create(h, dtype == MQR_F64, batch, m, n, lda, A, tau, tile, plan,
                               sched->per_cu, sched->order);
}

int mqr_batched_plan_run(mqr_handle_t h, mqr_batched_plan_t plan) {
    mqr::Range rg_("mqr_batched_plan_run");
    if (!plan) return -2;
    MQR_GUARD(h);
    if (batched_plan_device(plan) != h->device) return -2;
    MQR_BRIDGE(h);
    return batched_plan_run(h, plan);
}

int mqr_batched_plan_destroy(mqr_batched_plan_t plan) {
    if (!plan) return MQR_SUCCESS;
    DeviceGuard guard(batched_plan_device(plan));
    batched_plan_destroy(plan);
    return MQR_SUCCESS;
}

int mqr_qr_backward(mqr_handle_t h, mqr_dtype_t dtype, int64_t batch, int64_t m, int64_t n,
                    const void *Q, int64_t ldq, int64_t strideQ, const void *R, int64_t ldr,
                    int64_t strideR, const void *dQ, int64_t lddq, int64_t stridedQ,
                    const void *dR, int64_t lddr, int64_t stridedR, void *dA, int64_t ldda,
                    int64_t stridedA) {
    mqr::Range rg_("mqr_qr_backward");
    if (!h) return -1;
    if (dtype != MQR_F64 && dtype != MQR_F32) return -2;
    if (batch < 0 || batch > 65535) return -3;
    if (m < 0 || m > INT32_MAX) return -4;
    if (n < 0 || n > m) return -5;
    const bool any = batch > 0 && n > 0;
    if (any && !Q) return -6;
    if (ldq < std::max<int64_t>(1, m)) return -7;
    if (batch > 1 && strideQ < ldq * n) return -8;
    if (any && !R) return -9;
    if (ldr < std::max<int64_t>(1, n)) return -10;
    if (batch > 1 && strideR < ldr * n) return -11;
    if (dQ && lddq < std::max<int64_t>(1, m)) return -13;
    if (dQ && batch > 1 && stridedQ < lddq * n) return -14;
    if (dR && lddr < std::max<int64_t>(1, n)) return -16;
    if (dR && batch > 1 && stridedR < lddr * n) return -17;
    if (any && !dA) return -18;
    if (ldda < std::max<int64_t>(1, m)) return -19;
    if (batch > 1 && stridedA < ldda * n) return -20;
    if (!any) return MQR_SUCCESS;
    MQR_GUARD(h);
    MQR_BRIDGE(h);
    return qr_backward(h, dtype == MQR_F64, batch, m, n, Q, ldq, strideQ, R, ldr, strideR, dQ,
                       lddq, stridedQ, dR, lddr, stridedR, dA, ldda, stridedA);
}

int mqr_qr_householder_host(mqr_handle_t h, int64_t m, int64_t n, const double *A, double *Q,
                            double *R) {
    mqr::Range rg_("mqr_qr_householder_host");
    if (!h) return -1;
    if (m < 0) return -2;
    if (n < 0) return -3;
    if (m * n > 0 && !A) return -4;
    if (m > 0 && !Q) return -5;
    if (m * n > 0 && !R) return -6;
    if (m == 0) return MQR_SUCCESS;
    MQR_GUARD(h);
    MQR_BRIDGE(h);
    const int64_t k = std::min(m, n);
    const int nb = h->nb;
    double *dA = nullptr, *dQ = nullptr, *dTT = nullptr;
    int st = MQR_SUCCESS;
    const size_t szA = sizeof(double) * (size_t)(m * std::max<int64_t>(n, 1));
    const size_t szQ = sizeof(double) * (size_t)(m * m);
    const size_t szT = sizeof(double) * (size_t)(std::max<int64_t>(k, 1) * (nb + 1));
    if (hipMalloc(&dA, szA) != hipSuccess || hipMalloc(&dQ, szQ) != hipSuccess ||
        hipMalloc(&dTT, szT) != hipSuccess) {
        (void)hipGetLastError();
        st = MQR_ERR_ALLOC;
    }
    if (st == MQR_SUCCESS && n > 0 &&
        hipMemcpyAsync(dA, A, sizeof(double) * m * n, hipMemcpyHostToDevice, h->stream) != hipSuccess)
        st = MQR_ERR_HIP;
    double *dtau = dTT, *dT = dTT + std::max<int64_t>(k, 1);
    if (st == MQR_SUCCESS && k > 0) st = geqrf_f64(h, m, n, dA, m, dtau, dT, nb);
    if (st == MQR_SUCCESS) st = orgqr_f64(h, m, m, k, dA, m, dT, nb, dQ, m);
    if (st == MQR_SUCCESS &&
        (hipMemcpyAsync(Q, dQ, szQ, hipMemcpyDeviceToHost, h->stream) != hipSuccess ||
         (n > 0 && hipMemcpyAsync(R, dA, sizeof(double) * m * n, hipMemcpyDeviceToHost, h->stream) !=
                       hipSuccess) ||
         hipStreamSynchronize(h->stream) != hipSuccess))
        st = MQR_ERR_HIP;
    if (st == MQR_SUCCESS)
        for (int64_t c = 0; c < n; ++c)
            for (int64_t r = c + 1; r < m; ++r) R[r + c * m] = 0.0;  // QR/__Householder_Blocked.h:346-352
    if (dA) (void)hipFree(dA);
    if (dQ) (void)hipFree(dQ);
    if (dTT) (void)hipFree(dTT);
    return st;
}

int mqr_dgeqr_givens(mqr_handle_t h, int64_t m, int64_t n, const double *A, int64_t lda,
                     double *Q, int64_t ldq, double *R, int64_t ldr) {
    mqr::Range rg_("mqr_dgeqr_givens");
    if (!h) return -1;
    if (m < 0 || m > INT32_MAX) return -2;
    if (n < 0 || n > INT32_MAX) return -3;
    if (m * n > 0 && !A) return -4;
    if (lda < std::max<int64_t>(1, m)) return -5;
    if (m > 0 && !Q) return -6;
    if (ldq < std::max<int64_t>(1, m)) return -7;
    if (m * n > 0 && !R) return -8;
    if (ldr < std::max<int64_t>(1, m)) return -9;
    if (m == 0) return MQR_SUCCESS;
    if (!givens_supported(m, n)) return MQR_ERR_UNSUPPORTED;
    MQR_GUARD(h);
    MQR_BRIDGE(h);
    return qr_givens_f64(h, m, n, A, lda, Q, ldq, R, ldr);
}

int mqr_dgeqr_gramschmidt(mqr_handle_t h, int64_t m, int64_t n, const double *A, int64_t lda,
                          double *Q, int64_t ldq, double *R, int64_t ldr) {
    mqr::Range rg_("mqr_dgeqr_gramschmidt");
    if (!h) return -1;
    if (m < 0 || m > INT32_MAX) return -2;
    if (n < 0 || n > m) return -3;
    if (m * n > 0 && !A) return -4;
    if (lda < std::max<int64_t>(1, m)) return -5;
    if (m * n > 0 && !Q) return -6;
    if (ldq < std::max<int64_t>(1, m)) return -7;
    if (n > 0 && !R) return -8;
    if (ldr < std::max<int64_t>(1, n)) return -9;
    if (Q == A && ldq != lda) return -7;
    if (n == 0) return MQR_SUCCESS;
    MQR_GUARD(h);
    MQR_BRIDGE(h);
    return qr_gramschmidt_f64(h, m, n, A, lda, Q, ldq, R, ldr);
}

int mqr_qr_host(mqr_handle_t h, mqr_qr_type_t type, int64_t m, int64_t n, const double *A,
                double *Q, double *R) {
    mqr::Range rg_("mqr_qr_host");
    if (!h) return -1;
    if (type == MQR_QR_HOUSEHOLDER) {
        const int st = mqr_qr_householder_host(h, m, n, A, Q, R);
        return st < -1 ? st - 1 : st;  // argument positions shift by one
    }
    if (type != MQR_QR_GIVENS && type != MQR_QR_GRAMSCHMIDT) return -2;
    if (m < 0 || m > INT32_MAX) return -3;
    if (n < 0 || n > INT32_MAX || (type == MQR_QR_GRAMSCHMIDT && n > m)) return -4;
    if (m * n > 0 && !A) return -5;
    if (m > 0 && !Q) return -6;
    if (m * n > 0 && !R) return -7;
    if (m == 0) return MQR_SUCCESS;
    if (type == MQR_QR_GIVENS && !givens_supported(m, n)) return MQR_ERR_UNSUPPORTED;
    MQR_GUARD(h);
    MQR_BRIDGE(h);
    const int64_t qcols = type == MQR_QR_GIVENS ? m : n;
    double *d = nullptr;
    const size_t na = (size_t)(m * n), nq = (size_t)(m * qcols);
    if (hipMalloc(&d, sizeof(double) * std::max<size_t>(na + nq + na, 1)) != hipSuccess) {
        (void)hipGetLastError();
        return MQR_ERR_ALLOC;
    }
    double *dA = d, *dQ = d + na, *dR = d + na + nq;
    int st = MQR_SUCCESS;
    if (na && hipMemcpyAsync(dA, A, sizeof(double) * na, hipMemcpyHostToDevice, h->stream) !=
                  hipSuccess)
        st = MQR_ERR_HIP;
    if (st == MQR_SUCCESS) {
        if (type == MQR_QR_GIVENS) {
            st = qr_givens_f64(h, m, n, dA, m, dQ, m, dR, m);
        } else if (n > 0) {
            // R is m x n here: its rows n.. are zero, the n x n upper part
            // comes from the device routine (ld m)
            if (hipMemsetAsync(dR, 0, sizeof(double) * na, h->stream) != hipSuccess) st = MQR_ERR_HIP;
            if (st == MQR_SUCCESS) st = qr_gramschmidt_f64(h, m, n, dA, m, dQ, m, dR, m);
        }
    }
    if (st == MQR_SUCCESS &&
        ((nq && hipMemcpyAsync(Q, dQ, sizeof(double) * nq, hipMemcpyDeviceToHost, h->stream) !=
                    hipSuccess) ||
         (na && hipMemcpyAsync(R, dR, sizeof(double) * na, hipMemcpyDeviceToHost, h->stream) !=
                    hipSuccess) ||
         hipStreamSynchronize(h->stream) != hipSuccess))
        st = MQR_ERR_HIP;
    (void)hipStreamSynchronize(h->stream);
    (void)hipFree(d);
    return st;
}

// pivot and info scratch of an LU call (the caller's arrays when given)
static int lu_scratch(mqr_handle_t h, int64_t k, int32_t **piv, int32_t **info) {
    MQR_CHECK(ws_reserve(h, h->ws_lu, sizeof(int32_t) * (size_t)(k + 64)));
    *info = (int32_t *)h->ws_lu.ptr;
    *piv = *info + 16;
    return MQR_SUCCESS;
}

int mqr_dgetrf(mqr_handle_t h, int64_t m, int64_t n, double *A, int64_t lda, int32_t *ipiv,
               int32_t *info) {
    mqr::Range rg_("mqr_dgetrf");
    if (!h) return -1;
    if (m < 0 || m > INT32_MAX) return -2;
    if (n < 0 || n > INT32_MAX) return -3;
    if (m * n > 0 && !A) return -4;
    if (lda < std::max<int64_t>(1, m)) return -5;
    if (std::min(m, n) == 0) return MQR_SUCCESS;
    MQR_GUARD(h);
    MQR_BRIDGE(h);
    int32_t *sp = nullptr, *si = nullptr;
    MQR_CHECK(lu_scratch(h, std::min(m, n), &sp, &si));
    return getrf_f64(h, m, n, A, lda, ipiv, info, sp, si);
}

int mqr_dgetrs(mqr_handle_t h, int64_t n, int64_t nrhs, const double *LU, int64_t ldlu,
               const int32_t *ipiv, double *B, int64_t ldb) {
    mqr::Range rg_("mqr_dgetrs");
    if (!h) return -1;
    if (n < 0 || n > INT32_MAX) return -2;
    if (nrhs < 0 || nrhs > INT32_MAX) return -3;
    if (n > 0 && !LU) return -4;
    if (ldlu < std::max<int64_t>(1, n)) return -5;
    if (n * nrhs > 0 && !B) return -7;
    if (ldb < std::max<int64_t>(1, n)) return -8;
    if (n * nrhs == 0) return MQR_SUCCESS;
    MQR_GUARD(h);
    MQR_BRIDGE(h);
    return getrs_f64(h, n, nrhs, LU, ldlu, ipiv, B, ldb);
}

int mqr_dgesv(mqr_handle_t h, int64_t n, int64_t nrhs, double *A, int64_t lda, int32_t *ipiv,
              double *B, int64_t ldb, int32_t *info) {
    mqr::Range rg_("mqr_dgesv");
    if (!h) return -1;
    if (n < 0 || n > INT32_MAX) return -2;
    if (nrhs < 0 || nrhs > INT32_MAX) return -3;
    if (n > 0 && !A) return -4;
    if (lda < std::max<int64_t>(1, n)) return -5;
    if (n > 0 && !ipiv) return -6;
    if (n * nrhs > 0 && !B) return -7;
    if (ldb < std::max<int64_t>(1, n)) return -8;
    if (n == 0) return MQR_SUCCESS;
    MQR_GUARD(h);
    MQR_BRIDGE(h);
    int32_t *sp = nullptr, *si = nullptr;
    MQR_CHECK(lu_scratch(h, n, &sp, &si));
    MQR_CHECK(getrf_f64(h, n, n, A, lda, ipiv, info, sp, si));
    if (nrhs == 0) return MQR_SUCCESS;
    return getrs_f64(h, n, nrhs, A, lda, ipiv, B, ldb);
}

int mqr_lu_host(mqr_handle_t h, int pivot, int64_t m, int64_t n, const double *A, double *F,
                int32_t *pos, int32_t *info) {
    mqr::Range rg_("mqr_lu_host");
    if (!h) return -1;
    if (pivot != 0 && pivot != 1) return -2;
    if (m < 0 || m > INT32_MAX) return -3;
    if (n < 0 || n > INT32_MAX) return -4;
    if (m * n > 0 && !A) return -5;
    if (m * n > 0 && !F) return -6;
    if (m > 0 && !pos) return -7;
    for (int64_t i = 0; i < m; ++i) pos[i] = (int32_t)i;
    if (info) *info = 0;
    const int64_t k = std::min(m, n);
    if (k == 0) {
        if (m * n > 0) std::memcpy(F, A, sizeof(double) * (size_t)(m * n));
        return MQR_SUCCESS;
    }
    MQR_GUARD(h);
    MQR_BRIDGE(h);
    double *dA = nullptr;
    int32_t *dp = nullptr;
    const size_t na = (size_t)(m * n);
    if (hipMalloc(&dA, sizeof(double) * na) != hipSuccess ||
        hipMalloc(&dp, sizeof(int32_t) * (size_t)(k + 16)) != hipSuccess) {
        (void)hipGetLastError();
        if (dA) (void)hipFree(dA);
        return MQR_ERR_ALLOC;
    }
    int st = MQR_SUCCESS;
    std::vector<int32_t> ipiv((size_t)k);
    int32_t hinfo = 0;
    if (hipMemcpyAsync(dA, A, sizeof(double) * na, hipMemcpyHostToDevice, h->stream) != hipSuccess)
        st = MQR_ERR_HIP;
    if (st == MQR_SUCCESS) {
        int32_t *sp = nullptr, *si = nullptr;
        st = lu_scratch(h, k, &sp, &si);
        if (st == MQR_SUCCESS) st = getrf_f64(h, m, n, dA, m, pivot ? dp : nullptr, dp + k, sp, si);
    }
    if (st == MQR_SUCCESS &&
        (hipMemcpyAsync(F, dA, sizeof(double) * na, hipMemcpyDeviceToHost, h->stream) != hipSuccess ||
         (pivot && hipMemcpyAsync(ipiv.data(), dp, sizeof(int32_t) * (size_t)k,
                                  hipMemcpyDeviceToHost, h->stream) != hipSuccess) ||
         hipMemcpyAsync(&hinfo, dp + k, sizeof(int32_t), hipMemcpyDeviceToHost, h->stream) !=
             hipSuccess ||
         hipStreamSynchronize(h->stream) != hipSuccess))
        st = MQR_ERR_HIP;
    (void)hipStreamSynchronize(h->stream);
    (void)hipFree(dA);
    (void)hipFree(dp);
    if (st != MQR_SUCCESS) return st;
    // the sequence of swaps -> the row permutation pos (row i of L U is row
    // pos[i] of A), as the reference keeps it (LU/__PLU.h:37-40)
    if (pivot)
        for (int64_t i = 0; i < k; ++i) std::swap(pos[i], pos[ipiv[i]]);
    if (info) *info = hinfo;
    return MQR_SUCCESS;
}

int mqr_solve_direct_host(mqr_handle_t h, int64_t n, int64_t nrhs, const double *A,
                          const double *B, double *X) {
    mqr::Range rg_("mqr_solve_direct_host");
    if (!h) return -1;
    if (n < 0 || n > INT32_MAX) return -2;
    if (nrhs < 0 || nrhs > INT32_MAX) return -3;
    if (n > 0 && !A) return -4;
    if (n * nrhs > 0 && !B) return -5;
    if (n * nrhs > 0 && !X) return -6;
    if (n * nrhs == 0) return MQR_SUCCESS;
    MQR_GUARD(h);
    MQR_BRIDGE(h);
    double *d = nullptr;
    int32_t *dp = nullptr;
    const size_t na = (size_t)(n * n), nb = (size_t)(n * nrhs);
    if (hipMalloc(&d, sizeof(double) * (na + nb)) != hipSuccess ||
        hipMalloc(&dp, sizeof(int32_t) * (size_t)(n + 16)) != hipSuccess) {
        (void)hipGetLastError();
        if (d) (void)hipFree(d);
        return MQR_ERR_ALLOC;
    }
    int st = MQR_SUCCESS;
    if (hipMemcpyAsync(d, A, sizeof(double) * na, hipMemcpyHostToDevice, h->stream) != hipSuccess ||
        hipMemcpyAsync(d + na, B, sizeof(double) * nb, hipMemcpyHostToDevice, h->stream) !=
            hipSuccess)
        st = MQR_ERR_HIP;
    if (st == MQR_SUCCESS) {
        int32_t *sp = nullptr, *si = nullptr;
        st = lu_scratch(h, n, &sp, &si);
        if (st == MQR_SUCCESS) st = getrf_f64(h, n, n, d, n, dp, dp + n, sp, si);
        if (st == MQR_SUCCESS) st = getrs_f64(h, n, nrhs, d, n, dp, d + na, n);
    }
    if (st == MQR_SUCCESS &&
        (hipMemcpyAsync(X, d + na, sizeof(double) * nb, hipMemcpyDeviceToHost, h->stream) !=
             hipSuccess ||
         hipStreamSynchronize(h->stream) != hipSuccess))
        st = MQR_ERR_HIP;
    (void)hipStreamSynchronize(h->stream);
    (void)hipFree(d);
    (void)hipFree(dp);
    return st;
}

int mqr_gels_host(mqr_handle_t h, mqr_precision_t prec, int64_t m, int64_t n, const double *A,
                  const double *b, double *x, mqr_ir_info_t *info) {
    mqr::Range rg_("mqr_gels_host");
    if (!h) return -1;
    if (prec < MQR_PREC_FP64 || prec > MQR_PREC_BF16) return -2;
    if (m < 0) return -3;
    if (n < 0 || n > m) return -4;
    if (m * n > 0 && !A) return -5;
    if (m > 0 && !b) return -6;
    if (n > 0 && !x) return -7;
    if (info) std::memset(info, 0, sizeof(*info));
    if (n == 0) return MQR_SUCCESS;
    MQR_GUARD(h);
    MQR_BRIDGE(h);
    double *dA = nullptr, *db = nullptr, *dx = nullptr;
    int st = MQR_SUCCESS;
    if (hipMalloc(&dA, sizeof(double) * m * n) != hipSuccess ||
        hipMalloc(&db, sizeof(double) * m) != hipSuccess ||
        hipMalloc(&dx, sizeof(double) * std::max<int64_t>(n, 1)) != hipSuccess) {
        (void)hipGetLastError();
        st = MQR_ERR_ALLOC;
    }
    if (st == MQR_SUCCESS &&
        (hipMemcpyAsync(dA, A, sizeof(double) * m * n, hipMemcpyHostToDevice, h->stream) != hipSuccess ||
         hipMemcpyAsync(db, b, sizeof(double) * m, hipMemcpyHostToDevice, h->stream) != hipSuccess))
        st = MQR_ERR_HIP;
    if (st == MQR_SUCCESS) {
        if (prec == MQR_PREC_FP64 && !info) {
            st = mqr_dgels(h, m, n, 1, dA, m, db, m);
            if (st == MQR_SUCCESS &&
                hipMemcpyAsync(dx, db, sizeof(double) * n, hipMemcpyDeviceToDevice, h->stream) != hipSuccess)
                st = MQR_ERR_HIP;
        } else {
            st = gels_mixed(h, prec, m, n, dA, m, db, dx, nullptr, info);
        }
    }
    if (st == MQR_SUCCESS &&
        (hipMemcpyAsync(x, dx, sizeof(double) * n, hipMemcpyDeviceToHost, h->stream) != hipSuccess ||
         hipStreamSynchronize(h->stream) != hipSuccess))
        st = MQR_ERR_HIP;
    if (dA) (void)hipFree(dA);
    if (db) (void)hipFree(db);
    if (dx) (void)hipFree(dx);
    return st;
}

}  // extern "C"
