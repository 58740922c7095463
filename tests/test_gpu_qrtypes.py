"""GPU parity of the other QR_type values on the QR() switch
(QR/QR.h:26-42), through the C ABI.

- QR_Givens (QR/__Givens.h:45-63): the GPU follows the reference's rotation
  order and unfused arithmetic, so Q and R equal the oracle's restatement bit
  for bit on inputs without zero pivots; with zero pivots (triangular inputs)
  both take the documented limit rotation and stay exact.
- QR_GramSchmidt (QR/__GramSchmidt.h:2-18): the unique thin QR with a
  positive diagonal; the GPU's block Gram-Schmidt with reorthogonalisation
  agrees with the oracle's modified Gram-Schmidt to O(cond(A) u) and is
  orthogonal to working precision.
"""
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


@pytest.mark.parametrize("m,n", [(1, 1), (2, 1), (7, 4), (40, 40), (90, 33), (25, 60), (300, 200),
                                 (513, 129)])
def test_givens_bit_exact_against_oracle(h, orc, m, n):
    note(f"givens {m}x{n}")
    rng = np.random.default_rng(m * 13 + n)
    A = rng.standard_normal((m, n))
    Q, R = M.qr_givens(h, M.colmajor(A))
    h.synchronize()
    Q, R = host(Q), host(R)
    Qo, Ro = orc.qr_givens(A, fix_zero_pivots=True)
    dq, dr = np.abs(Q - Qo).max(), np.abs(R - Ro).max()
    note(f"givens {m}x{n}: max |Q - Qo| {dq:.3g}, |R - Ro| {dr:.3g}")
    assert np.array_equal(Q, Qo) and np.array_equal(R, Ro)
    nA = np.linalg.norm(A)
    assert np.linalg.norm(Q @ R - A) / nA < 16 * max(m, n) * EPS


def test_givens_zero_pivots_and_host_api(h, orc):
    A = np.triu(np.random.default_rng(2).standard_normal((9, 6)))
    Q, R = M.qr_givens(h, M.colmajor(A))
    h.synchronize()
    assert np.array_equal(host(R), A) and np.array_equal(host(Q), np.eye(9))
    B = np.array([[0.0, 1.0], [-3.0, 2.0]])
    Qb, Rb = M.qr_givens(h, M.colmajor(B))
    Qo, Ro = orc.qr_givens(B, fix_zero_pivots=True)
    assert np.array_equal(host(Qb), Qo) and np.array_equal(host(Rb), Ro)
    # the host call behind QR(A, Q, R, QR_Givens)
    A = np.asfortranarray(np.random.default_rng(3).standard_normal((50, 20)))
    Qh = np.zeros((50, 50), order="F")
    Rh = np.zeros((50, 20), order="F")
    M.QR(A, Qh, Rh, M.QR_Givens)
    Qo, Ro = orc.qr_givens(A, fix_zero_pivots=True)
    assert np.array_equal(Qh, Qo) and np.array_equal(Rh, Ro)


def test_givens_argument_errors(h):
    L = M.lib()
    p = C.c_void_p(1)
    assert L.mqr_dgeqr_givens(h.h, -1, 2, p, 1, p, 1, p, 1) == -2
    assert L.mqr_dgeqr_givens(h.h, 4, 2, p, 3, p, 4, p, 4) == -5
    assert L.mqr_dgeqr_givens(h.h, 4, 2, p, 4, p, 4, p, 3) == -9
    # one sweep step would hold more rotations than the kernel's table
    assert L.mqr_dgeqr_givens(h.h, 20000, 9000, p, 20000, p, 20000, p, 20000) == \
        M.MQR_ERR_UNSUPPORTED


def haar_np(rng, rows, cols):
    Q, R = np.linalg.qr(rng.standard_normal((rows, cols)))
    return Q * np.sign(np.diag(R))[None, :]


@pytest.mark.parametrize("m,n,cond", [(1, 1, 1), (7, 4, 1), (100, 64, 1), (300, 200, 1),
                                      (1000, 130, 1), (513, 257, 1), (800, 300, 1e6)])
def test_gramschmidt_against_oracle(h, orc, m, n, cond):
    note(f"gram-schmidt {m}x{n} cond {cond:g}")
    rng = np.random.default_rng(m + 7 * n)
    if cond == 1:
        A = rng.standard_normal((m, n))
    else:
        A = (haar_np(rng, m, n) * np.logspace(0, -np.log10(cond), n)) @ haar_np(rng, n, n).T
    Q, R = M.qr_gramschmidt(h, M.colmajor(A))
    h.synchronize()
    Q, R = host(Q), host(R)
    nA = np.linalg.norm(A)
    k = np.linalg.cond(A)
    assert np.all(np.tril(R, -1) == 0.0) and np.all(np.diag(R) > 0.0)
    assert np.linalg.norm(Q @ R - A) / nA < 16 * m * EPS
    orth = np.linalg.norm(Q.T @ Q - np.eye(n))
    assert orth < 16 * m * EPS
    Qo, Ro = orc.qr_gramschmidt(A)
    dr = np.abs(R - Ro[:n]).max() / nA
    dq = np.abs(Q - Qo).max()
    note(f"gram-schmidt {m}x{n}: |QtQ-I| {orth:.2e}, |R-Ro|/|A| {dr:.2e}, |Q-Qo| {dq:.2e}, "
         f"cond {k:.2e}")
    assert dr < 64 * m * EPS * k and dq < 64 * m * EPS * k


def test_gramschmidt_in_place_host_api_and_errors(h, orc):
    rng = np.random.default_rng(8)
    A = rng.standard_normal((200, 90))
    dA = M.colmajor(A)
    R = M.empty_colmajor(90, 90)
    L = M.lib()
    M.check(L.mqr_dgeqr_gramschmidt(h.h, 200, 90, M._ptr(dA), 200, M._ptr(dA), 200, M._ptr(R),
                                    90), "in place")
    h.synchronize()
    Q2, R2 = M.qr_gramschmidt(h, M.colmajor(A))
    h.synchronize()
    assert np.array_equal(host(dA), host(Q2)) and np.array_equal(host(R), host(R2))
    Qh = np.zeros((200, 90), order="F")
    Rh = np.zeros((200, 90), order="F")
    M.QR(np.asfortranarray(A), Qh, Rh, M.QR_GramSchmidt)
    assert np.array_equal(Qh, host(Q2)) and np.array_equal(Rh[:90], host(R2))
    assert np.all(Rh[90:] == 0.0)
    p = C.c_void_p(1)
    assert L.mqr_dgeqr_gramschmidt(h.h, 3, 4, p, 3, p, 3, p, 4) == -3
    assert L.mqr_dgeqr_gramschmidt(h.h, 4, 3, p, 4, p, 4, p, 2) == -9
