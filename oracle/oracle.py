"""ctypes front end of the CPU oracle (TEST INFRASTRUCTURE, NOT PRODUCT CODE).

The shared libraries under ``oracle/lib`` hold a C restatement of Mat-8.3.1's
Householder QR path (``oracle/mqr_oracle.c``; each function there cites the
reference lines it follows).  Only ``tests/``, ``__graft_entry__.smoke()`` and
``bench.py``'s cpu_baseline leg import this module, and only as the checker or
the timed CPU baseline.  The ``mat_amd`` product never imports it.

All matrices are numpy float64 arrays in Fortran (column-major) order, like
``MAT<T>`` (Base/__Matrix.h:141-146).
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_DIR = os.path.join(_HERE, "lib")
_libs: dict[str, C.CDLL] = {}

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="F_CONTIGUOUS")


def build() -> None:
    """Compile both oracle variants with gcc (make -C oracle)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def cpu_has_avx512() -> bool:
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("flags"):
                    fl = line.split()
                    return "avx512f" in fl and "avx512dq" in fl
    except OSError:
        pass
    return False


def _bind(lib: C.CDLL) -> C.CDLL:
    i64 = C.c_int64
    vp = C.c_void_p
    lib.orc_config.argtypes = [vp]
    lib.orc_work_doubles.argtypes = [C.c_int, C.c_int]
    lib.orc_work_doubles.restype = C.c_size_t
    lib.orc_dot.argtypes = [vp, vp, C.c_int]
    lib.orc_dot.restype = C.c_double
    lib.orc_pack_A.argtypes = [vp, i64, C.c_int, C.c_int, vp]
    lib.orc_pack_B.argtypes = [vp, i64, C.c_int, C.c_int, vp]
    lib.orc_micro_kernel.argtypes = [vp, vp, vp, i64, C.c_int, C.c_int, C.c_int, C.c_int]
    lib.orc_mulAB_packed_n.argtypes = [vp, vp, vp, i64, C.c_int, C.c_int, C.c_int]
    lib.orc_gemm.argtypes = [C.c_int, C.c_int, C.c_int, vp, i64, vp, i64, vp, i64]
    lib.orc_panel.argtypes = [C.c_int, C.c_int, vp, i64, vp, vp]
    lib.orc_larft.argtypes = [C.c_int, C.c_int, vp, i64, vp, vp, vp, vp]
    lib.orc_larfb_t.argtypes = [C.c_int, C.c_int, C.c_int, vp, i64, vp, vp, i64, vp]
    lib.orc_larfb_n.argtypes = [C.c_int, C.c_int, C.c_int, vp, i64, vp, vp, i64, vp]
    lib.orc_geqrf_blocked.argtypes = [C.c_int, C.c_int, vp, vp, vp, vp]
    lib.orc_orgqr_full.argtypes = [C.c_int, C.c_int, vp, vp, vp, vp]
    lib.orc_qr_unblocked.argtypes = [C.c_int, C.c_int, vp, vp]
    lib.orc_qr_householder.argtypes = [C.c_int, C.c_int, vp, vp, vp]
    lib.orc_qr_householder.restype = C.c_int
    lib.orc_gels.argtypes = [C.c_int, C.c_int, vp, vp, vp]
    lib.orc_qr_givens.argtypes = [C.c_int, C.c_int, vp, vp, vp, C.c_int]
    lib.orc_qr_gramschmidt.argtypes = [C.c_int, C.c_int, vp, vp, vp]
    lib.orc_lu_factor.argtypes = [C.c_int, C.c_int, vp, C.c_int, vp]
    lib.orc_lu_factor.restype = C.c_int
    lib.orc_solve_direct.argtypes = [C.c_int, C.c_int, vp, vp, vp]
    lib.orc_trsv_upper.argtypes = [C.c_int, vp, i64, vp]
    return lib


def resolve(variant: str) -> str:
    """``"auto"`` -> the host's widest ISA; ``"auto_simd"`` likewise with the
    register-accumulator micro-kernel."""
    isa = "avx512" if cpu_has_avx512() else "avx2"
    return {"auto": isa, "auto_simd": isa + "_simd"}.get(variant, variant)


def load(variant: str = "avx2") -> C.CDLL:
    """Load one oracle build: ``"avx2"`` (BS 96, 8x6 tile) or ``"avx512"``
    (BS 48, 16x12 tile), ``"avx2_simd"`` / ``"avx512_simd"`` (the same with
    the micro-kernel's accumulators in vector registers, bit-identical), or
    ``"auto"`` / ``"auto_simd"`` (AVX-512 when the host has it)."""
    variant = resolve(variant)
    if variant not in _libs:
        path = os.path.join(_LIB_DIR, f"liborc_{variant}.so")
        if not os.path.exists(path):
            build()
        _libs[variant] = _bind(C.CDLL(path))
    return _libs[variant]


def _f(a) -> np.ndarray:
    return np.asfortranarray(np.asarray(a, dtype=np.float64))


def _p(a: np.ndarray) -> int:
    return a.ctypes.data


class Oracle:
    """One oracle build with numpy-level helpers."""

    def __init__(self, variant: str = "avx2"):
        self.variant = resolve(variant)
        self.lib = load(self.variant)
        cfg = (C.c_int * 6)()
        self.lib.orc_config(C.cast(cfg, C.c_void_p))
        self.MR, self.NR, self.MC, self.KC, self.NC, self.BS = list(cfg)

    # -- sizes -------------------------------------------------------------
    def work(self, m: int, n: int) -> np.ndarray:
        return np.zeros(int(self.lib.orc_work_doubles(m, n)) + 8, dtype=np.float64)

    # -- level 1 / GEMM layer ------------------------------------------------
    def dot(self, a, b) -> float:
        a = np.ascontiguousarray(a, dtype=np.float64)
        b = np.ascontiguousarray(b, dtype=np.float64)
        return float(self.lib.orc_dot(_p(a), _p(b), len(a)))

    def gemm(self, A, B) -> np.ndarray:
        A, B = _f(A), _f(B)
        M, K = A.shape
        K2, N = B.shape
        assert K == K2
        Cm = np.zeros((M, N), dtype=np.float64, order="F")
        self.lib.orc_gemm(M, N, K, _p(A), max(M, 1), _p(B), max(K, 1), _p(Cm), max(M, 1))
        return Cm

    def pack_A(self, A) -> np.ndarray:
        A = _f(A)
        mc, kc = A.shape
        out = np.full(((mc + self.MR - 1) // self.MR) * kc * self.MR, np.nan)
        self.lib.orc_pack_A(_p(A), mc, mc, kc, _p(out))
        return out

    def pack_B(self, B) -> np.ndarray:
        B = _f(B)
        kc, nc = B.shape
        out = np.full(((nc + self.NR - 1) // self.NR) * kc * self.NR, np.nan)
        self.lib.orc_pack_B(_p(B), kc, kc, nc, _p(out))
        return out

    def mulAB_packed_n(self, A, B, Cm) -> np.ndarray:
        """C - A*B through the pack routines and the nmul macro-kernel."""
        A, B = _f(A), _f(B)
        Cm = _f(Cm).copy(order="F")
        m, k = A.shape
        n = B.shape[1]
        Ap, Bp = self.pack_A(A), self.pack_B(B)
        self.lib.orc_mulAB_packed_n(_p(Ap), _p(Bp), _p(Cm), m, m, n, k)
        return Cm

    # -- QR pieces ------------------------------------------------------------
    def panel(self, P):
        P = _f(P).copy(order="F")
        h, b = P.shape
        tau = np.zeros(b)
        v = np.zeros(h + 1)
        self.lib.orc_panel(h, b, _p(P), h, _p(tau), _p(v))
        return P, tau

    def larft(self, V, tau) -> np.ndarray:
        V = _f(V)
        h, b = V.shape
        tau = np.ascontiguousarray(tau, dtype=np.float64)
        T = np.zeros((b, b), order="F")
        s = np.zeros(b + 1)
        w = np.zeros(b + 1)
        self.lib.orc_larft(h, b, _p(V), h, _p(tau), _p(T), _p(s), _p(w))
        return T

    def larfb(self, V, T, Cm, transposed: bool = True) -> np.ndarray:
        V, T = _f(V), _f(T)
        Cm = _f(Cm).copy(order="F")
        h, b = V.shape
        nc = Cm.shape[1]
        work = self.work(h, max(h, nc))
        fn = self.lib.orc_larfb_t if transposed else self.lib.orc_larfb_n
        fn(h, b, nc, _p(V), h, _p(T), _p(Cm), h, _p(work))
        return Cm

    def geqrf_blocked(self, A):
        """Returns (VR, T_all, tau): R on/above the diagonal, reflectors
        below, and the packed T blocks (block j at T_all[j*BS:], ld = b_j)."""
        R = _f(A).copy(order="F")
        m, n = R.shape
        k = min(m, n)
        T_all = np.zeros(((k + self.BS - 1) // self.BS) * self.BS * self.BS + 1)
        tau = np.zeros(max(k, 1))
        work = self.work(m, n)
        self.lib.orc_geqrf_blocked(m, n, _p(R), _p(T_all), _p(tau), _p(work))
        return R, T_all, tau[:k]

    def t_block(self, T_all, j: int, k: int) -> np.ndarray:
        """The b x b T factor of the block whose first column is j."""
        b = min(self.BS, k - j)
        return np.array(T_all[j * self.BS: j * self.BS + b * b]).reshape((b, b), order="F")

    def orgqr_full(self, VR, T_all, n: int) -> np.ndarray:
        VR = _f(VR)
        m = VR.shape[0]
        Q = np.zeros((m, m), order="F")
        T_all = np.ascontiguousarray(T_all, dtype=np.float64)
        work = self.work(m, n)
        self.lib.orc_orgqr_full(m, n, _p(VR), _p(T_all), _p(Q), _p(work))
        return Q

    def qr_unblocked(self, A):
        R = _f(A).copy(order="F")
        m, n = R.shape
        Q = np.zeros((m, m), order="F")
        self.lib.orc_qr_unblocked(m, n, _p(Q), _p(R))
        return Q, R

    def qr(self, A):
        """QR(A, Q, R, QR_Householder): returns (Q, R, blocked_branch)."""
        A = _f(A)
        m, n = A.shape
        Q = np.zeros((m, m), order="F")
        R = np.zeros((m, n), order="F")
        blocked = self.lib.orc_qr_householder(m, n, _p(A), _p(Q), _p(R))
        return Q, R, bool(blocked)

    def qr_givens(self, A, fix_zero_pivots: bool = False):
        """QR(A, Q, R, QR_Givens): returns (Q m x m, R m x n)."""
        A = _f(A)
        m, n = A.shape
        Q = np.zeros((m, m), order="F")
        R = np.zeros((m, n), order="F")
        self.lib.orc_qr_givens(m, n, _p(A), _p(Q), _p(R), 1 if fix_zero_pivots else 0)
        return Q, R

    def qr_gramschmidt(self, A):
        """QR(A, Q, R, QR_GramSchmidt): returns (Q m x n, R m x n)."""
        A = _f(A)
        m, n = A.shape
        Q = np.zeros((m, n), order="F")
        R = np.zeros((m, n), order="F")
        self.lib.orc_qr_gramschmidt(m, n, _p(A), _p(Q), _p(R))
        return Q, R

    def lu(self, A, pivot: bool = True):
        """LU / PLU in place on a copy: returns (F, pos, factored) with L
        unit lower below F's diagonal and U on and above, row i of L U being
        row pos[i] of A."""
        F = _f(A).copy(order="F")
        m, n = F.shape
        pos = np.zeros(max(m, 1), dtype=np.int32)
        done = self.lib.orc_lu_factor(m, n, _p(F), 1 if pivot else 0, pos.ctypes.data)
        return F, pos[:m], int(done)

    def solve_direct(self, A, B) -> np.ndarray:
        A, B = _f(A), _f(B)
        n = A.shape[0]
        nrhs = B.shape[1]
        X = np.zeros((n, nrhs), order="F")
        self.lib.orc_solve_direct(n, nrhs, _p(A), _p(B), _p(X))
        return X

    def gels(self, A, b) -> np.ndarray:
        A = _f(A)
        m, n = A.shape
        b = np.ascontiguousarray(b, dtype=np.float64)
        x = np.zeros(n)
        self.lib.orc_gels(m, n, _p(A), _p(b), _p(x))
        return x

    def trsv_upper(self, R, c) -> np.ndarray:
        R = _f(R)
        x = np.array(c, dtype=np.float64)
        self.lib.orc_trsv_upper(R.shape[1], _p(R), R.shape[0], _p(x))
        return x


# ---------------------------------------------------------------------------
# numpy restatements used by the checks of the new forms (no reference
# counterpart): QR adjoint and least-squares backward error.
# ---------------------------------------------------------------------------

def qr_adjoint(Q: np.ndarray, R: np.ndarray, dQ: np.ndarray, dR: np.ndarray) -> np.ndarray:
    """dA = [dQ + Q copyltu(M)] R^{-T}, M = R dR^T - dQ^T Q, for the thin QR
    of an m x n (m >= n) matrix; copyltu(M) = tril(M) + tril(M, -1)^T."""
    M = R @ dR.T - dQ.T @ Q
    cl = np.tril(M) + np.tril(M, -1).T
    B = dQ + Q @ cl
    # solve X R^T = B  <=>  R X^T = B^T
    import scipy.linalg as sla
    return sla.solve_triangular(R, B.T, lower=False).T


def ls_backward_error(A: np.ndarray, b: np.ndarray, x: np.ndarray) -> dict:
    """Normwise backward errors of an approximate least-squares solution.

    ``kw``: Karlson-Walden estimate ||(A^T A + phi^2 I)^{-1/2} A^T r|| / ||x||
    with phi = ||r|| / ||x||, divided by ||A||_F; ``cheap``:
    ||A^T r|| / (||A||_F (||A||_F ||x|| + ||b||))."""
    r = b - A @ x
    nx = np.linalg.norm(x)
    nr = np.linalg.norm(r)
    nA = np.linalg.norm(A)
    Atr = A.T @ r
    cheap = np.linalg.norm(Atr) / (nA * (nA * nx + np.linalg.norm(b)))
    phi = nr / nx if nx > 0 else 0.0
    U, s, Vt = np.linalg.svd(A, full_matrices=False)
    y = Vt @ Atr
    kw = np.linalg.norm(y / np.sqrt(s * s + phi * phi)) / nx / nA if nx > 0 else 0.0
    return {"kw": float(kw), "cheap": float(cheap), "resid": float(nr)}
