"""Python mirror of the MI355X Householder-QR library's C ABI (include/mat_qr.h).

The product is ``mat_amd/lib/libmatqr.so`` (HIP, gfx950).  This module binds
its ``extern "C"`` entry points with ctypes and adds thin helpers for
column-major device buffers held in torch tensors (torch supplies device
memory and streams only).  There is no CPU fallback: every call goes to the
HIP library, and loading fails loudly when the library is missing.

Column-major convention: a logical m x n matrix is a torch tensor of shape
(m, n) with strides (1, ld), as returned by :func:`colmajor`; element (r, c)
sits at ``data[r + c * ld]`` like ``MAT<T>`` (Base/__Matrix.h:141-146).
"""
from __future__ import annotations

import ctypes as C
import os
from typing import Optional

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libmatqr.so")

# QR_type of the reference (Mat.h:43-47); only QR_Householder is on the GPU path.
QR_Householder, QR_Givens, QR_GramSchmidt = 0, 1, 2

# status codes (include/mat_qr.h)
MQR_SUCCESS = 0
MQR_ERR_NO_DEVICE = 1001
MQR_ERR_HIP = 1002
MQR_ERR_ALLOC = 1003
MQR_ERR_NOT_CONVERGED = 1004
MQR_ERR_UNSUPPORTED = 1005
MQR_ERR_NOT_INITIALIZED = 1006
MQR_ERR_INTERNAL = 1007
MQR_ERR_TOO_LARGE = 1008
MQR_ERR_IO = 1009
NOALLOC_MAX_BATCH = 32768

MQR_F64, MQR_F32 = 0, 1
PREC = {"fp64": 0, "fp32": 1, "bf16x3": 2, "bf16": 3}

# Every exported symbol of include/mat_qr.h with its ctypes signature.
_i64, _i32, _vp, _dbl = C.c_int64, C.c_int, C.c_void_p, C.c_double
SIGNATURES = {
    "mqr_strerror": (C.c_char_p, [_i32]),
    "mqr_get_device_count": (_i32, [_vp]),
    "mqr_create": (_i32, [_vp, _i32]),
    "mqr_destroy": (_i32, [_vp]),
    "mqr_set_stream": (_i32, [_vp, _vp]),
    "mqr_get_stream": (_i32, [_vp, _vp]),
    "mqr_set_block_size": (_i32, [_vp, _i32]),
    "mqr_get_block_size": (_i32, [_vp, _vp]),
    "mqr_synchronize": (_i32, [_vp]),
    "mqr_malloc": (_i32, [_vp, _vp, C.c_size_t]),
    "mqr_free": (_i32, [_vp, _vp]),
    "mqr_memcpy_htod": (_i32, [_vp, _vp, _vp, C.c_size_t]),
    "mqr_memcpy_dtoh": (_i32, [_vp, _vp, _vp, C.c_size_t]),
    "mqr_profile_enable": (_i32, [_vp, _i32]),
    "mqr_profile_read": (_i32, [_vp, _vp, _i32, _vp]),
    "mqr_profile_set_mask": (_i32, [_vp, C.c_uint]),
    "mqr_dgemm": (_i32, [_vp, C.c_char, C.c_char, _i64, _i64, _i64, _dbl, _vp, _i64, _vp, _i64,
                         _dbl, _vp, _i64]),
    "mqr_sgemm": (_i32, [_vp, C.c_char, C.c_char, _i64, _i64, _i64, C.c_float, _vp, _i64, _vp,
                         _i64, C.c_float, _vp, _i64]),
    "mqr_sgemm_bf16": (_i32, [_vp, _i32, C.c_char, C.c_char, _i64, _i64, _i64, C.c_float, _vp,
                              _i64, _vp, _i64, C.c_float, _vp, _i64]),
    "mqr_dgeqrf": (_i32, [_vp, _i64, _i64, _vp, _i64, _vp, _vp, _i64]),
    "mqr_dorgqr": (_i32, [_vp, _i64, _i64, _i64, _vp, _i64, _vp, _i64, _vp, _i64]),
    "mqr_dormqr": (_i32, [_vp, C.c_char, _i64, _i64, _i64, _vp, _i64, _vp, _i64, _vp, _i64]),
    "mqr_dtrsm_upper": (_i32, [_vp, C.c_char, _i64, _i64, _vp, _i64, _vp, _i64]),
    "mqr_dgels": (_i32, [_vp, _i64, _i64, _i64, _vp, _i64, _vp, _i64]),
    "mqr_gels_mixed": (_i32, [_vp, _i32, _i64, _i64, _vp, _i64, _vp, _vp, _vp, _vp]),
    "mqr_geqrf_batched_uneven": (_i32, [_vp, _i32, _i64, _vp, _vp, _vp, _vp, _vp, _i32]),
    "mqr_geqrf_batched_uneven_noalloc": (_i32, [_vp, _i32, _i64, _vp, _vp, _vp, _vp, _vp, _i32]),
    "mqr_batched_plan_create": (_i32, [_vp, _i32, _i64, _vp, _vp, _vp, _vp, _vp, _i32, _vp]),
    "mqr_batched_plan_create_sched": (_i32, [_vp, _i32, _i64, _vp, _vp, _vp, _vp, _vp, _vp,
                                             _vp]),
    "mqr_batched_plan_run": (_i32, [_vp, _vp]),
    "mqr_batched_plan_destroy": (_i32, [_vp]),
    "mqr_qr_backward": (_i32, [_vp, _i32, _i64, _i64, _i64, _vp, _i64, _i64, _vp, _i64, _i64,
                               _vp, _i64, _i64, _vp, _i64, _i64, _vp, _i64, _i64]),
    "mqr_dgemm_host": (_i32, [_vp, _i64, _i64, _i64, _vp, _vp, _vp]),
    "mqr_sgemm_host": (_i32, [_vp, _i64, _i64, _i64, _vp, _vp, _vp]),
    "mqr_file_size": (_i32, [C.c_char_p, _vp]),
    "mqr_mat_save": (_i32, [C.c_char_p, C.c_int32, C.c_int32, _vp, C.c_size_t]),
    "mqr_mat_load_header": (_i32, [C.c_char_p, _vp, _vp]),
    "mqr_mat_load": (_i32, [C.c_char_p, C.c_int32, C.c_int32, _vp, C.c_size_t]),
    "mqr_raw_save": (_i32, [C.c_char_p, _vp, C.c_size_t]),
    "mqr_raw_load": (_i32, [C.c_char_p, _vp, C.c_size_t]),
    "mqr_dgeqr_givens": (_i32, [_vp, _i64, _i64, _vp, _i64, _vp, _i64, _vp, _i64]),
    "mqr_dgeqr_gramschmidt": (_i32, [_vp, _i64, _i64, _vp, _i64, _vp, _i64, _vp, _i64]),
    "mqr_qr_host": (_i32, [_vp, _i32, _i64, _i64, _vp, _vp, _vp]),
    "mqr_dgetrf": (_i32, [_vp, _i64, _i64, _vp, _i64, _vp, _vp]),
    "mqr_dgetrs": (_i32, [_vp, _i64, _i64, _vp, _i64, _vp, _vp, _i64]),
    "mqr_dgesv": (_i32, [_vp, _i64, _i64, _vp, _i64, _vp, _vp, _i64, _vp]),
    "mqr_lu_host": (_i32, [_vp, _i32, _i64, _i64, _vp, _vp, _vp, _vp]),
    "mqr_solve_direct_host": (_i32, [_vp, _i64, _i64, _vp, _vp, _vp]),
    "mqr_qr_householder_host": (_i32, [_vp, _i64, _i64, _vp, _vp, _vp]),
    "mqr_gels_host": (_i32, [_vp, _i32, _i64, _i64, _vp, _vp, _vp, _vp]),
}


class IROpts(C.Structure):
    _fields_ = [("max_iter", C.c_int), ("tol", C.c_double), ("fallback", C.c_int),
                ("krylov_max_iter", C.c_int)]


class IRInfo(C.Structure):
    _fields_ = [("iterations", C.c_int), ("converged", C.c_int), ("fell_back", C.c_int),
                ("backward_error", C.c_double), ("residual_norm", C.c_double),
                ("krylov_iterations", C.c_int)]


class KernelStat(C.Structure):
    _fields_ = [("name", C.c_char * 24), ("launches", C.c_int64), ("total_ms", C.c_double),
                ("flops", C.c_double), ("bytes", C.c_double)]


# kernel kinds of the profiler, in the order of their mask bit (mqr_kernel_kind_t)
KINDS = ("panel_gemm", "vta_gemm", "panel_small", "w2", "update_gemm", "trsm", "gemm", "resid",
         "cast", "other", "lookahead", "apply_q", "near_update")


def profile_enable(h: "Handle", on: bool = True, kinds=None) -> None:
    """Turn HIP-event timing on or off; `kinds` restricts it to some kinds."""
    mask = 0xFFFFFFFF if kinds is None else sum(1 << KINDS.index(k) for k in kinds)
    check(lib().mqr_profile_set_mask(h.h, mask), "mqr_profile_set_mask")
    check(lib().mqr_profile_enable(h.h, 1 if on else 0), "mqr_profile_enable")


def profile_read(h: "Handle") -> dict:
    """{kind: {launches, total_ms, flops, bytes}} since the last read/enable."""
    arr = (KernelStat * 16)()
    cnt = C.c_int(0)
    check(lib().mqr_profile_read(h.h, C.cast(arr, C.c_void_p), 16, C.byref(cnt)),
          "mqr_profile_read")
    out = {}
    for i in range(min(cnt.value, 16)):
        s = arr[i]
        out[s.name.decode()] = {"launches": int(s.launches), "total_ms": float(s.total_ms),
                                "flops": float(s.flops), "bytes": float(s.bytes)}
    return out


class MqrError(RuntimeError):
    def __init__(self, status: int, where: str):
        self.status = status
        super().__init__(f"{where} failed with status {status}: {strerror(status)}")


_lib: Optional[C.CDLL] = None


def lib() -> C.CDLL:
    """Load libmatqr.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it with `python -m mat_amd.build`")
        # torch ships its own HIP runtime under the same SONAME
        # (libamdhip64.so.7).  Loading torch first makes libmatqr bind to that
        # already-loaded runtime, so one process never holds two of them.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def strerror(status: int) -> str:
    return lib().mqr_strerror(status).decode()


def check(status: int, where: str) -> None:
    if status != MQR_SUCCESS:
        raise MqrError(status, where)


def device_count() -> int:
    n = C.c_int(0)
    lib().mqr_get_device_count(C.byref(n))
    return n.value


class Handle:
    """mqr_handle_t: device, stream (torch's current stream by default) and
    workspace."""

    def __init__(self, device: int = 0, use_torch_stream: bool = True, nb: Optional[int] = None):
        self.h = C.c_void_p()
        check(lib().mqr_create(C.byref(self.h), device), "mqr_create")
        self.device = device
        if use_torch_stream:
            import torch
            s = torch.cuda.current_stream(device)
            check(lib().mqr_set_stream(self.h, C.c_void_p(s.cuda_stream)), "mqr_set_stream")
        if nb is not None:
            check(lib().mqr_set_block_size(self.h, nb), "mqr_set_block_size")

    @property
    def nb(self) -> int:
        v = C.c_int(0)
        check(lib().mqr_get_block_size(self.h, C.byref(v)), "mqr_get_block_size")
        return v.value

    def synchronize(self) -> None:
        check(lib().mqr_synchronize(self.h), "mqr_synchronize")

    def close(self) -> None:
        if self.h:
            check(lib().mqr_destroy(self.h), "mqr_destroy")
            self.h = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ---------------------------------------------------------------------------
# column-major device buffers in torch tensors
# ---------------------------------------------------------------------------

def colmajor(x, device="cuda", dtype=None):
    """A new (m, n) torch tensor with strides (1, m) holding a copy of x
    (numpy or torch); never a view of x."""
    import numpy as np
    import torch
    if isinstance(x, np.ndarray):
        t = torch.from_numpy(np.ascontiguousarray(x.T))
    else:
        t = x.T.clone(memory_format=torch.contiguous_format)
    if dtype is not None:
        t = t.to(dtype)
    return t.to(device).T


def empty_colmajor(m: int, n: int, device="cuda", dtype=None):
    import torch
    return torch.empty((n, m), device=device, dtype=dtype or torch.float64).T


def ld(t) -> int:
    """Leading dimension of a column-major (m, n) tensor (strides of size-1
    dimensions are irrelevant and not checked)."""
    assert t.dim() == 2, "expected a matrix"
    m, n = t.shape
    if n <= 1 or t.numel() == 0:
        assert t.numel() <= 1 or n == 0 or t.stride(0) == 1, "expected a column-major tensor"
        return max(m, 1)
    assert t.stride(0) == 1 or m <= 1, "expected a column-major tensor"
    return max(t.stride(1), m, 1)


def _ptr(t) -> C.c_void_p:
    return C.c_void_p(t.data_ptr() if t is not None else 0)


# ---------------------------------------------------------------------------
# device-level entry points
# ---------------------------------------------------------------------------

def gemm_bf16(h: Handle, terms: int, transa: str, transb: str, alpha: float, A, B, beta: float,
              Cm):
    """C = alpha op(A) op(B) + beta C for float32 column-major tensors on the
    BF16 MFMA (terms 1: hi*hi; 3: hi*hi + hi*lo + lo*hi), FP32 accumulation."""
    m, n = Cm.shape
    k = A.shape[0] if transa == "T" else A.shape[1]
    check(lib().mqr_sgemm_bf16(h.h, terms, transa.encode(), transb.encode(), m, n, k, alpha,
                               _ptr(A), ld(A), _ptr(B), ld(B), beta, _ptr(Cm), ld(Cm)),
          "mqr_sgemm_bf16")
    return Cm


def gemm(h: Handle, transa: str, transb: str, alpha: float, A, B, beta: float, Cm):
    """C = alpha op(A) op(B) + beta C on column-major device tensors (float64
    through mqr_dgemm, float32 through mqr_sgemm)."""
    import torch
    m, n = Cm.shape
    k = A.shape[0] if transa == "T" else A.shape[1]
    if Cm.dtype == torch.float32:
        check(lib().mqr_sgemm(h.h, transa.encode(), transb.encode(), m, n, k, alpha, _ptr(A),
                              ld(A), _ptr(B), ld(B), beta, _ptr(Cm), ld(Cm)), "mqr_sgemm")
        return Cm
    check(lib().mqr_dgemm(h.h, transa.encode(), transb.encode(), m, n, k, alpha, _ptr(A), ld(A),
                          _ptr(B), ld(B), beta, _ptr(Cm), ld(Cm)), "mqr_dgemm")
    return Cm


def geqrf(h: Handle, A):
    """Factor the column-major float64 device matrix A in place; returns
    (tau, T) with T of shape (nb, k) in the dgeqrt layout."""
    import torch
    m, n = A.shape
    k = min(m, n)
    tau = torch.zeros(max(k, 1), device=A.device, dtype=torch.float64)
    nb = h.nb
    T = empty_colmajor(nb, max(k, 1), device=A.device)
    T.zero_()
    check(lib().mqr_dgeqrf(h.h, m, n, _ptr(A), ld(A), _ptr(tau), _ptr(T), nb), "mqr_dgeqrf")
    return tau[:k], T


def orgqr(h: Handle, VR, T, k: int, nq: Optional[int] = None):
    m = VR.shape[0]
    nq = m if nq is None else nq
    Q = empty_colmajor(m, nq, device=VR.device)
    check(lib().mqr_dorgqr(h.h, m, nq, k, _ptr(VR), ld(VR), _ptr(T), ld(T), _ptr(Q), ld(Q)),
          "mqr_dorgqr")
    return Q


def ormqr(h: Handle, trans: str, VR, T, k: int, Cm):
    m, nrhs = Cm.shape
    check(lib().mqr_dormqr(h.h, trans.encode(), m, nrhs, k, _ptr(VR), ld(VR), _ptr(T), ld(T),
                           _ptr(Cm), ld(Cm)), "mqr_dormqr")
    return Cm


def trsm_upper(h: Handle, trans: str, R, B):
    n = R.shape[1]
    check(lib().mqr_dtrsm_upper(h.h, trans.encode(), n, B.shape[1], _ptr(R), ld(R), _ptr(B), ld(B)),
          "mqr_dtrsm_upper")
    return B


def gels(h: Handle, A, B):
    """Plain FP64 least squares; A and B are overwritten, returns X = B[:n]."""
    m, n = A.shape
    check(lib().mqr_dgels(h.h, m, n, B.shape[1], _ptr(A), ld(A), _ptr(B), ld(B)), "mqr_dgels")
    return B[:n]


def gels_mixed(h: Handle, prec: str, A, b, max_iter: int = 30, tol: float = 0.0,
               fallback: bool = True, krylov_max_iter: int = 0):
    """Mixed-precision least squares with refinement; returns (x, IRInfo)."""
    import torch
    m, n = A.shape
    x = torch.zeros(n, device=A.device, dtype=torch.float64)
    opts = IROpts(max_iter, tol, 1 if fallback else 0, krylov_max_iter)
    info = IRInfo()
    check(lib().mqr_gels_mixed(h.h, PREC[prec], m, n, _ptr(A), ld(A), _ptr(b), _ptr(x),
                               C.byref(opts), C.byref(info)), "mqr_gels_mixed")
    return x, info


def _batch_args(mats, taus, device: int):
    import numpy as np
    import torch
    b = len(mats)
    f64 = mats[0].dtype == torch.float64 if b else True
    # the kernel indexes every buffer with the batch's one element type: a
    # mixed list, a matrix on another device or a row-major view would be
    # read and written out of bounds
    for i, (t, tau) in enumerate(zip(mats, taus)):
        if t.dtype not in (torch.float64, torch.float32) or t.dtype != mats[0].dtype:
            raise TypeError(f"matrix {i}: dtype {t.dtype}, expected {mats[0].dtype} like matrix 0")
        if tau.dtype != t.dtype:
            raise TypeError(f"tau {i}: dtype {tau.dtype}, expected {t.dtype}")
        if (t.device.type != "cuda" or t.device.index != device or tau.device != t.device):
            raise ValueError(f"matrix {i}: on {t.device} (tau on {tau.device}), expected the "
                             f"handle's device cuda:{device}")
        if t.dim() != 2 or (t.shape[0] > 1 and t.shape[1] > 1 and t.stride(0) != 1):
            raise ValueError(f"matrix {i}: expected a column-major (m, n) tensor")
        if tau.numel() < min(t.shape):
            raise ValueError(f"tau {i}: {tau.numel()} entries, expected at least {min(t.shape)}")
    m = np.array([t.shape[0] for t in mats], dtype=np.int32)
    n = np.array([t.shape[1] for t in mats], dtype=np.int32)
    lda = np.array([ld(t) for t in mats], dtype=np.int64)
    A = (C.c_void_p * max(b, 1))(*[t.data_ptr() for t in mats])
    T = (C.c_void_p * max(b, 1))(*[t.data_ptr() for t in taus])
    keep = (m, n, lda, A, T)
    return (0 if f64 else 1), b, keep


def geqrf_batched(h: Handle, mats, tile: int = -1):
    """Factor a list of column-major device matrices (all float64 or all
    float32, m <= 512) in place in one launch; returns their tau vectors."""
    import torch
    taus = [torch.zeros(max(min(t.shape), 1), device=t.device, dtype=t.dtype) for t in mats]
    if not mats:
        return taus
    dt, b, (m, n, lda, A, T) = _batch_args(mats, taus, h.device)
    check(lib().mqr_geqrf_batched_uneven(h.h, dt, b, m.ctypes.data_as(C.c_void_p),
                                         n.ctypes.data_as(C.c_void_p),
                                         lda.ctypes.data_as(C.c_void_p), C.cast(A, C.c_void_p),
                                         C.cast(T, C.c_void_p), tile), "mqr_geqrf_batched_uneven")
    return taus


def geqrf_batched_noalloc(h: Handle, mats, taus, tile: int = -1):
    """The same factorisation through mqr_geqrf_batched_uneven_noalloc: the
    caller supplies the tau vectors (nothing is allocated on the device)."""
    if not mats:
        return taus
    dt, b, (m, n, lda, A, T) = _batch_args(mats, taus, h.device)
    check(lib().mqr_geqrf_batched_uneven_noalloc(h.h, dt, b, m.ctypes.data_as(C.c_void_p),
                                                 n.ctypes.data_as(C.c_void_p),
                                                 lda.ctypes.data_as(C.c_void_p),
                                                 C.cast(A, C.c_void_p), C.cast(T, C.c_void_p),
                                                 tile), "mqr_geqrf_batched_uneven_noalloc")
    return taus


class BatchedPlan:
    """The batched factorisation of fixed device matrices captured in a
    hipGraph; run() replays it on the handle's stream."""

    def __init__(self, h: Handle, mats, tile: int = -1, per_cu: Optional[int] = None,
                 order: Optional[int] = None):
        """tile: -1 tuned, 0 panel width 8, 1 width 16.  per_cu / order: the
        launch's schedule (mqr_batched_schedule_t); None takes the tuned one."""
        import torch
        self.h = h
        self.mats = mats
        self.taus = [torch.zeros(max(min(t.shape), 1), device=t.device, dtype=t.dtype)
                     for t in mats]
        self.plan = C.c_void_p()
        dt, b, (m, n, lda, A, T) = _batch_args(mats, self.taus, h.device)
        args = (m.ctypes.data_as(C.c_void_p), n.ctypes.data_as(C.c_void_p),
                lda.ctypes.data_as(C.c_void_p), C.cast(A, C.c_void_p), C.cast(T, C.c_void_p))
        if per_cu is None and order is None:
            check(lib().mqr_batched_plan_create(h.h, dt, b, *args, tile, C.byref(self.plan)),
                  "mqr_batched_plan_create")
        else:
            sched = (C.c_int * 3)({-1: 0, 0: 8, 1: 16}[tile],
                                  -1 if per_cu is None else per_cu, -1 if order is None else order)
            check(lib().mqr_batched_plan_create_sched(h.h, dt, b, *args, C.cast(sched, C.c_void_p),
                                                      C.byref(self.plan)),
                  "mqr_batched_plan_create_sched")

    def run(self):
        check(lib().mqr_batched_plan_run(self.h.h, self.plan), "mqr_batched_plan_run")

    def close(self):
        if self.plan:
            lib().mqr_batched_plan_destroy(self.plan)
            self.plan = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def colmajor_batched(x):
    """(batch, m, n) tensor with each matrix column-major (strides (m n, 1, m))."""
    return x.transpose(1, 2).contiguous().transpose(1, 2)


def _bstride(t) -> int:
    return t.stride(0) if t.shape[0] > 1 else t.shape[1] * t.shape[2]


def _bld(t) -> int:
    assert t.dim() == 3 and (t.stride(1) == 1 or t.shape[1] <= 1), "expected column-major matrices"
    return max(t.stride(2) if t.shape[2] > 1 else t.shape[1], t.shape[1], 1)


def qr_backward(h: Handle, Q, R, dQ=None, dR=None):
    """dA for a batch of thin QRs A = Q R (tensors (batch, m, n) / (batch, n,
    n), column-major per matrix, float32 or float64) via mqr_qr_backward."""
    import torch
    bsz, m, n = Q.shape
    dA = colmajor_batched(torch.empty((bsz, m, n), device=Q.device, dtype=Q.dtype))
    dt = 0 if Q.dtype == torch.float64 else 1

    def arg(t):
        return (_ptr(t), _bld(t), _bstride(t)) if t is not None else (_ptr(None), 1, 0)

    check(lib().mqr_qr_backward(h.h, dt, bsz, m, n, *arg(Q), *arg(R), *arg(dQ), *arg(dR),
                                *arg(dA)), "mqr_qr_backward")
    return dA


def qr_thin_batched(h: Handle, A):
    """Thin FP64 QR of each matrix of a (batch, m, n) tensor on the GPU
    (mqr_dgeqrf + mqr_dorgqr): returns Q (batch, m, n), R (batch, n, n)."""
    import torch
    bsz, m, n = A.shape
    Qs, Rs = [], []
    for i in range(bsz):
        VR = colmajor(A[i].detach(), device=A.device)
        _, T = geqrf(h, VR)
        Q = orgqr(h, VR, T, n, n)
        Qs.append(Q)
        Rs.append(torch.triu(VR[:n, :n]))
    return colmajor_batched(torch.stack(Qs)), colmajor_batched(torch.stack(Rs))


_torch_handle: Optional[Handle] = None


def torch_handle() -> Handle:
    """A handle ordered on torch's current CUDA stream (re-read at every call)."""
    global _torch_handle
    import torch
    if _torch_handle is None:
        _torch_handle = Handle(0)
    s = torch.cuda.current_stream(0)
    check(lib().mqr_set_stream(_torch_handle.h, C.c_void_p(s.cuda_stream)), "mqr_set_stream")
    return _torch_handle


def _qr_function():
    import torch

    class QRFunction(torch.autograd.Function):
        """Thin QR with the library's forward (FP64 blocked Householder) and
        backward (mqr_qr_backward).  Test harness for gradient checks."""

        @staticmethod
        def forward(ctx, A):
            h = torch_handle()
            Q, R = qr_thin_batched(h, A)
            ctx.save_for_backward(Q, R)
            return Q, R

        @staticmethod
        def backward(ctx, gQ, gR):
            Q, R = ctx.saved_tensors
            h = torch_handle()
            gQ = colmajor_batched(gQ) if gQ is not None else None
            gR = colmajor_batched(gR) if gR is not None else None
            return qr_backward(h, Q, R, gQ, gR)

    return QRFunction


_QRF = None


def qr_autograd(A):
    """Q, R = thin QR of a (batch, m, n) float64 CUDA tensor, differentiable
    through mqr_qr_backward."""
    global _QRF
    if _QRF is None:
        _QRF = _qr_function()
    return _QRF.apply(A)


# ---------------------------------------------------------------------------
# host-level calls with the reference's signatures
# ---------------------------------------------------------------------------

_default_handle: Optional[Handle] = None


def default_handle() -> Handle:
    global _default_handle
    if _default_handle is None:
        _default_handle = Handle(0, use_torch_stream=False)
    return _default_handle


def QR(A, Q, R, qr_type: int = QR_Householder, handle: Optional[Handle] = None) -> None:
    """QR(A, Q, R, type) on host float64 Fortran-ordered numpy arrays, like
    QR/QR.h:7-46, every type on the GPU: Q (m x m; m x n for QR_GramSchmidt)
    and R (m x n) are written in place."""
    import numpy as np
    if qr_type not in (QR_Householder, QR_Givens, QR_GramSchmidt):
        raise MqrError(-2, "QR")
    h = handle or default_handle()
    A = np.asfortranarray(A, dtype=np.float64)
    m, n = A.shape
    qc = n if qr_type == QR_GramSchmidt else m
    assert Q.shape == (m, qc) and R.shape == (m, n)
    assert Q.flags.f_contiguous and R.flags.f_contiguous
    assert Q.dtype == np.float64 and R.dtype == np.float64
    check(lib().mqr_qr_host(h.h, qr_type, m, n, A.ctypes.data_as(C.c_void_p),
                            Q.ctypes.data_as(C.c_void_p), R.ctypes.data_as(C.c_void_p)),
          "mqr_qr_host")


def qr_givens(h: Handle, A):
    """QR by Givens rotations of a column-major float64 device matrix:
    returns new (Q m x m, R m x n) device tensors (mqr_dgeqr_givens)."""
    m, n = A.shape
    Q = empty_colmajor(m, m, device=A.device)
    R = empty_colmajor(m, n, device=A.device)
    check(lib().mqr_dgeqr_givens(h.h, m, n, _ptr(A), ld(A), _ptr(Q), ld(Q), _ptr(R), ld(R)),
          "mqr_dgeqr_givens")
    return Q, R


def qr_gramschmidt(h: Handle, A):
    """Gram-Schmidt thin QR (positive diagonal) of a column-major float64
    device matrix, m >= n: returns new (Q m x n, R n x n) device tensors
    (mqr_dgeqr_gramschmidt)."""
    m, n = A.shape
    Q = empty_colmajor(m, n, device=A.device)
    R = empty_colmajor(n, n, device=A.device)
    check(lib().mqr_dgeqr_gramschmidt(h.h, m, n, _ptr(A), ld(A), _ptr(Q), ld(Q), _ptr(R),
                                      ld(R)), "mqr_dgeqr_gramschmidt")
    return Q, R


def LeastSquares(A, b, prec: str = "fp64", handle: Optional[Handle] = None):
    """x = argmin ||A x - b|| on host arrays; returns (x, IRInfo)."""
    import numpy as np
    h = handle or default_handle()
    A = np.asfortranarray(A, dtype=np.float64)
    b = np.ascontiguousarray(b, dtype=np.float64)
    m, n = A.shape
    x = np.zeros(n)
    info = IRInfo()
    check(lib().mqr_gels_host(h.h, PREC[prec], m, n, A.ctypes.data_as(C.c_void_p),
                              b.ctypes.data_as(C.c_void_p), x.ctypes.data_as(C.c_void_p),
                              C.byref(info) if prec != "fp64" else None), "mqr_gels_host")
    return x, info


# ---------------------------------------------------------------------------
# binary MAT files (Mat-5.2/_extra/__File.h:1-66) through the library
# ---------------------------------------------------------------------------

def mat_save(path: str, A) -> None:
    """Write a 2-D numpy array as a MAT file: int32 rows, int32 cols, then the
    elements column-major (the storage order of MAT<T>), no type tag."""
    import numpy as np
    A = np.asfortranarray(A)
    assert A.ndim == 2
    check(lib().mqr_mat_save(os.fsencode(path), A.shape[0], A.shape[1],
                             A.ctypes.data_as(C.c_void_p), A.itemsize), "mqr_mat_save")


def mat_load(path: str, dtype="float64"):
    """Read a MAT file written with elements of `dtype` (the file carries
    no type tag); returns a Fortran-ordered numpy array."""
    import numpy as np
    r, c = C.c_int32(0), C.c_int32(0)
    check(lib().mqr_mat_load_header(os.fsencode(path), C.byref(r), C.byref(c)),
          "mqr_mat_load_header")
    A = np.empty((r.value, c.value), dtype=dtype, order="F")
    check(lib().mqr_mat_load(os.fsencode(path), r.value, c.value, A.ctypes.data_as(C.c_void_p),
                             A.itemsize), "mqr_mat_load")
    return A


def gemm_host(A, B):
    """C = A @ B for host float64 or float32 numpy matrices through the GPU
    GEMM behind MAT<T>::operator* (mqr_dgemm_host / mqr_sgemm_host)."""
    import numpy as np
    f32 = A.dtype == np.float32 and B.dtype == np.float32
    dt = np.float32 if f32 else np.float64
    A = np.asfortranarray(A, dtype=dt)
    B = np.asfortranarray(B, dtype=dt)
    assert A.shape[1] == B.shape[0]
    Cm = np.zeros((A.shape[0], B.shape[1]), dtype=dt, order="F")
    fn = lib().mqr_sgemm_host if f32 else lib().mqr_dgemm_host
    check(fn(default_handle().h, A.shape[0], B.shape[1], A.shape[1], A.ctypes.data_as(C.c_void_p),
             B.ctypes.data_as(C.c_void_p), Cm.ctypes.data_as(C.c_void_p)),
          "mqr_sgemm_host" if f32 else "mqr_dgemm_host")
    return Cm


# ---------------------------------------------------------------------------
# LU, PLU and the direct solve
# ---------------------------------------------------------------------------

def getrf(h: Handle, A, pivot: bool = True):
    """LU of a column-major float64 device matrix in place (L \\ U); returns
    (ipiv device int32 tensor or None, info device int32 tensor)."""
    import torch
    m, n = A.shape
    k = min(m, n)
    ipiv = torch.zeros(max(k, 1), device=A.device, dtype=torch.int32) if pivot else None
    info = torch.zeros(1, device=A.device, dtype=torch.int32)
    check(lib().mqr_dgetrf(h.h, m, n, _ptr(A), ld(A), _ptr(ipiv), _ptr(info)), "mqr_dgetrf")
    return ipiv, info


def getrs(h: Handle, LU, ipiv, B):
    n = LU.shape[0]
    check(lib().mqr_dgetrs(h.h, n, B.shape[1], _ptr(LU), ld(LU), _ptr(ipiv), _ptr(B), ld(B)),
          "mqr_dgetrs")
    return B


def LU(A, pivot: bool = True, handle: Optional[Handle] = None):
    """LU(A, L, U) / PLU(A, pos, L, U) of a host float64 matrix on the GPU:
    returns (F, pos, info) with F = L \\ U (m x n, Fortran order)."""
    import numpy as np
    h = handle or default_handle()
    A = np.asfortranarray(A, dtype=np.float64)
    m, n = A.shape
    F = np.zeros((m, n), order="F")
    pos = np.zeros(max(m, 1), dtype=np.int32)
    info = C.c_int32(0)
    check(lib().mqr_lu_host(h.h, 1 if pivot else 0, m, n, A.ctypes.data_as(C.c_void_p),
                            F.ctypes.data_as(C.c_void_p), pos.ctypes.data_as(C.c_void_p),
                            C.byref(info)), "mqr_lu_host")
    return F, pos[:m], info.value


def Solve_direct(A, B, handle: Optional[Handle] = None):
    """X = A^{-1} B (the reference's Solve_direct) for host float64 arrays."""
    import numpy as np
    h = handle or default_handle()
    A = np.asfortranarray(A, dtype=np.float64)
    B = np.asfortranarray(B, dtype=np.float64)
    if B.ndim == 1:
        B = B[:, None]
    n, nrhs = B.shape
    X = np.zeros((n, nrhs), order="F")
    check(lib().mqr_solve_direct_host(h.h, n, nrhs, A.ctypes.data_as(C.c_void_p),
                                      B.ctypes.data_as(C.c_void_p), X.ctypes.data_as(C.c_void_p)),
          "mqr_solve_direct_host")
    return X
