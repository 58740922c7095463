"""GPU parity of LU, PLU and Solve_direct (LU/__LU.h, LU/__PLU.h,
LU/__Solve_direct.h) through the C ABI, against the oracle's restatement
(itself pinned to LAPACK dgetrf: the same first-largest pivot rule).  The
permutation must be identical, and L and U are equal bit for bit: every
element takes the same fused multiply-adds in the same order (the panel's
rank-1 updates, the triangular solve and the MFMA GEMM's k-ordered
accumulation all subtract l_ik u_kj for k = 0, 1, ... as the oracle's
left-looking column updates do)."""
import ctypes as C

import numpy as np
import pytest

import mat_amd as M
from oracle.oracle import Oracle

pytestmark = pytest.mark.gpu
EPS = np.finfo(np.float64).eps


def note(msg):
    print(msg, flush=True)


@pytest.fixture(scope="module")
def torch_mod():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a GPU"
    return torch


@pytest.fixture(scope="module")
def h(torch_mod):
    hd = M.Handle(0)
    yield hd
    hd.close()


@pytest.fixture(scope="module")
def orc():
    return Oracle("avx2")


def host(t):
    return np.asfortranarray(t.detach().cpu().numpy())


def perm_from_ipiv(ipiv, m):
    pos = np.arange(m)
    for i, p in enumerate(ipiv):
        pos[[i, p]] = pos[[p, i]]
    return pos


@pytest.mark.parametrize("m,n", [(1, 1), (5, 5), (33, 33), (100, 64), (64, 100), (300, 300),
                                 (513, 257), (1000, 1000), (2000, 2000)])
def test_plu_matches_oracle(h, orc, m, n):
    note(f"plu {m}x{n}")
    rng = np.random.default_rng(m * 3 + n)
    A = rng.standard_normal((m, n))
    dA = M.colmajor(A)
    ipiv, info = M.getrf(h, dA)
    h.synchronize()
    F = host(dA)
    k = min(m, n)
    pos = perm_from_ipiv(ipiv.cpu().numpy()[:k], m)
    Fo, poso, done = orc.lu(A)
    assert int(info.cpu()[0]) == 0 and done == k
    assert np.array_equal(pos, poso)
    err = np.abs(F - Fo).max()
    note(f"plu {m}x{n}: max |F - F_oracle| {err:.2e}")
    assert np.array_equal(F, Fo)
    L = np.tril(F[:, :k], -1) + np.eye(m, k)
    U = np.triu(F[:k])
    assert np.linalg.norm(L @ U - A[pos]) <= 16 * max(m, n) * EPS * np.linalg.norm(A)


def test_lu_without_pivoting(h, orc):
    rng = np.random.default_rng(4)
    A = rng.standard_normal((300, 300)) + 300 * np.eye(300)
    dA = M.colmajor(A)
    ipiv, info = M.getrf(h, dA, pivot=False)
    h.synchronize()
    Fo, pos, done = orc.lu(A, pivot=False)
    assert ipiv is None and int(info.cpu()[0]) == 0 and done == 300
    assert np.array_equal(host(dA), Fo)


def test_zero_pivot_is_reported(h):
    rng = np.random.default_rng(5)
    A = rng.standard_normal((100, 100))
    A[:, 40] = 0.0                      # a zero column: no pivot for column 40
    dA = M.colmajor(A)
    _, info = M.getrf(h, dA)
    h.synchronize()
    assert int(info.cpu()[0]) == 41


@pytest.mark.parametrize("n,nrhs", [(1, 1), (7, 3), (100, 1), (250, 6), (1500, 40)])
def test_solve_direct_matches_oracle(h, orc, n, nrhs):
    note(f"solve {n} nrhs={nrhs}")
    rng = np.random.default_rng(n + 7 * nrhs)
    A = rng.standard_normal((n, n))
    B = rng.standard_normal((n, nrhs))
    dA, dB = M.colmajor(A), M.colmajor(B)
    ipiv = __import__("torch").zeros(n, device="cuda", dtype=__import__("torch").int32)
    M.check(M.lib().mqr_dgesv(h.h, n, nrhs, M._ptr(dA), n, M._ptr(ipiv), M._ptr(dB), n,
                              None), "mqr_dgesv")
    h.synchronize()
    X = host(dB)
    Xo = orc.solve_direct(A, B)
    cond = np.linalg.cond(A)
    np.testing.assert_allclose(X, Xo, rtol=0, atol=64 * n * EPS * cond * np.abs(Xo).max())
    # the host call behind Solve_direct(A, B)
    Xh = M.Solve_direct(A, B)
    assert np.array_equal(Xh, X)


def test_host_lu_api(orc):
    rng = np.random.default_rng(9)
    A = rng.standard_normal((120, 90))
    F, pos, info = M.LU(A)
    Fo, poso, _ = orc.lu(A)
    assert info == 0 and np.array_equal(pos, poso)
    np.testing.assert_allclose(F, Fo, rtol=0, atol=64 * 120 * EPS * np.abs(A).max())


def test_lu_argument_errors(h):
    L = M.lib()
    p = C.c_void_p(1)
    assert L.mqr_dgetrf(h.h, -1, 2, p, 1, p, p) == -2
    assert L.mqr_dgetrf(h.h, 4, 2, p, 3, p, p) == -5
    assert L.mqr_dgetrs(h.h, 4, 1, p, 3, p, p, 4) == -5
    assert L.mqr_dgesv(h.h, 4, 1, p, 4, None, p, 4, None) == -6
