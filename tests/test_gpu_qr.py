"""GPU parity of the FP64 path: every call goes through the C ABI
(libmatqr.so) and is compared with the CPU oracle (a restatement of the
reference's blocked Householder QR), LAPACK vectors, brute-force products and
size-independent invariants.

Tolerances: the reference's results are FP64 with O(n u) rounding; the GPU
accumulates in a different order (MFMA tiles, split-K partials), so elementwise
agreement is asked to c * max(m, n) * u * ||A|| with c = 64 (R, V, tau, Q) and
exact equality where all arithmetic is exact (integer-valued GEMM inputs,
identity input)."""
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
    """Column-major device tensor -> Fortran-ordered numpy array."""
    return np.asfortranarray(t.detach().cpu().numpy())


def unit_lower(VR, b):
    V = np.tril(VR[:, :b], -1)
    V[np.arange(b), np.arange(b)] = 1.0
    return V


def gpu_factor(h, A):
    dA = M.colmajor(A)
    tau, T = M.geqrf(h, dA)
    h.synchronize()
    return dA, tau, T


# ---------------------------------------------------------------------------
# MFMA GEMM tile
# ---------------------------------------------------------------------------

@pytest.mark.parametrize("ta,tb", [("N", "N"), ("T", "N"), ("N", "T"), ("T", "T")])
def test_gemm_exact_on_integers(h, orc, ta, tb):
    rng = np.random.default_rng(1)
    for (m, n, k) in [(1, 1, 1), (16, 16, 4), (64, 64, 16), (37, 29, 41), (200, 150, 90),
                      (1000, 70, 300), (513, 384, 200), (128, 128, 16), (260, 1030, 33)]:
        note(f"gemm {ta}{tb} m={m} n={n} k={k}")
        A = rng.integers(-8, 9, (k, m) if ta == "T" else (m, k)).astype(np.float64)
        B = rng.integers(-8, 9, (n, k) if tb == "T" else (k, n)).astype(np.float64)
        Cm = rng.integers(-8, 9, (m, n)).astype(np.float64)
        opA = A.T if ta == "T" else A
        opB = B.T if tb == "T" else B
        exact = Cm.astype(np.int64) - opA.astype(np.int64) @ opB.astype(np.int64)
        dC = M.colmajor(Cm)
        M.gemm(h, ta, tb, -1.0, M.colmajor(A), M.colmajor(B), 1.0, dC)
        got = host(dC)
        assert np.array_equal(got, exact.astype(np.float64)), (m, n, k)
        if ta == "N" and tb == "N":
            assert np.array_equal(got, orc.mulAB_packed_n(opA, opB, Cm))


def test_gemm_float_against_oracle(h, orc):
    rng = np.random.default_rng(2)
    A = rng.standard_normal((300, 200))
    B = rng.standard_normal((200, 170))
    dC = M.colmajor(np.zeros((300, 170)))
    M.gemm(h, "N", "N", 1.0, M.colmajor(A), M.colmajor(B), 0.0, dC)
    np.testing.assert_allclose(host(dC), orc.gemm(A, B), rtol=0, atol=200 * 4 * EPS * 20)


# ---------------------------------------------------------------------------
# factorisation parity
# ---------------------------------------------------------------------------

def test_geqrf_matches_lapack_vectors(h, orc, golden):
    nb = h.nb
    for key in golden.files:
        if not key.startswith("A_"):
            continue
        m, n = (int(x) for x in key[2:].split("x"))
        note(f"geqrf fixture {m}x{n}")
        A = golden[key]
        dA, tau, T = gpu_factor(h, A)
        VR, tau, T = host(dA), tau.cpu().numpy(), host(T)
        VRl, taul = golden[f"VR_{m}x{n}"], golden[f"tau_{m}x{n}"]
        k = min(m, n)
        tol = 64 * max(m, n) * EPS * max(1.0, np.abs(A).max())
        kk = k - 1 if m <= n else k
        np.testing.assert_allclose(tau[:kk], taul[:kk], rtol=0, atol=tol)
        np.testing.assert_allclose(np.triu(VR)[:kk], np.triu(VRl)[:kk], rtol=0, atol=tol)
        np.testing.assert_allclose(np.tril(VR, -1), np.tril(VRl, -1), rtol=0, atol=tol)
        if m <= n:
            assert tau[k - 1] == 2.0
            np.testing.assert_allclose(VR[k - 1, k - 1:], -VRl[k - 1, k - 1:], rtol=0, atol=tol)
        # every T block against the oracle's recurrence on the same reflectors
        for j in range(0, k, nb):
            b = min(nb, k - j)
            V = unit_lower(VR[j:, j:j + b], b)
            Tref = orc.larft(V, tau[j:j + b])
            np.testing.assert_allclose(np.triu(T[:b, j:j + b]), Tref, rtol=0,
                                       atol=64 * m * EPS * max(1.0, np.abs(Tref).max()))


@pytest.mark.parametrize("m,n", [(100, 64), (300, 250), (500, 500), (1024, 1024), (777, 333),
                                 (250, 400), (2048, 96), (65, 130)])
def test_qr_matches_oracle(h, orc, m, n):
    note(f"qr vs oracle {m}x{n}")
    rng = np.random.default_rng(m * 7 + n)
    A = rng.standard_normal((m, n))
    dA, tau, T = gpu_factor(h, A)
    k = min(m, n)
    dQ = M.orgqr(h, dA, T, k, nq=m)
    h.synchronize()
    VR, Q = host(dA), host(dQ)
    Qo, Ro, _ = orc.qr(A)
    nA = np.linalg.norm(A)
    tol = 64 * max(m, n) * EPS
    np.testing.assert_allclose(np.triu(VR), Ro, rtol=0, atol=tol * np.abs(A).max() * 4)
    np.testing.assert_allclose(Q, Qo, rtol=0, atol=tol)
    R = np.triu(VR)
    assert np.linalg.norm(Q @ R - A) / nA < 16 * max(m, n) * EPS
    assert np.linalg.norm(Q.T @ Q - np.eye(m)) < 16 * m * EPS * np.sqrt(m)


def test_thin_q_and_apply(h, orc):
    rng = np.random.default_rng(5)
    m, n = 900, 300
    A = rng.standard_normal((m, n))
    dA, tau, T = gpu_factor(h, A)
    dQt = M.orgqr(h, dA, T, n, nq=n)
    dQf = M.orgqr(h, dA, T, n, nq=m)
    Cm = rng.standard_normal((m, 5))
    dC = M.colmajor(Cm)
    M.ormqr(h, "T", dA, T, n, dC)
    QtC = host(dC)
    M.ormqr(h, "N", dA, T, n, dC)
    back = host(dC)
    h.synchronize()
    Qf = host(dQf)
    np.testing.assert_allclose(host(dQt), Qf[:, :n], rtol=0, atol=0)
    np.testing.assert_allclose(QtC, Qf.T @ Cm, rtol=0, atol=1e-12)
    np.testing.assert_allclose(back, Cm, rtol=0, atol=1e-12)


def test_trsm_upper(h):
    rng = np.random.default_rng(6)
    n = 333
    R = np.triu(rng.standard_normal((n, n))) + 4 * np.eye(n) * np.sign(rng.standard_normal(n))
    B = rng.standard_normal((n, 3))
    import scipy.linalg as sla
    for trans in ["N", "T"]:
        dB = M.colmajor(B)
        M.trsm_upper(h, trans, M.colmajor(R), dB)
        ref = sla.solve_triangular(R, B, trans=0 if trans == "N" else 1, lower=False)
        # random triangular R is ill-conditioned: bound relative to the solution size
        np.testing.assert_allclose(host(dB), ref, rtol=0, atol=1e-12 * n * np.abs(ref).max())


@pytest.mark.parametrize("n", [1, 63, 64, 65, 333, 1100])
def test_trsv_upper_one_rhs(h, n):
    """One right-hand side takes the per-block back substitution."""
    rng = np.random.default_rng(60 + n)
    R = np.triu(rng.standard_normal((n, n))) + 4 * np.eye(n) * np.sign(rng.standard_normal(n))
    b = rng.standard_normal((n, 1))
    import scipy.linalg as sla
    db = M.colmajor(b)
    M.trsm_upper(h, "N", M.colmajor(R), db)
    ref = sla.solve_triangular(R, b, lower=False)
    np.testing.assert_allclose(host(db), ref, rtol=0, atol=1e-12 * n * np.abs(ref).max())


@pytest.mark.parametrize("nrhs", [1, 2])
def test_dgels_qtb_contract(h, nrhs):
    """mqr_dgels leaves Q^T B in B: rows [0, n) hold X and the norm of rows
    [n, m) is the residual norm ||B - A X|| (Q orthogonal)."""
    rng = np.random.default_rng(70 + nrhs)
    m, n = 2000, 640
    A = rng.standard_normal((m, n))
    B = rng.standard_normal((m, nrhs))
    dA, dB = M.colmajor(A), M.colmajor(B)
    X = host(M.gels(h, dA, dB))
    QtB = host(dB)
    Xref = np.linalg.lstsq(A, B, rcond=None)[0]
    np.testing.assert_allclose(X, Xref, rtol=0, atol=1e-10 * max(1, np.abs(Xref).max()))
    for c in range(nrhs):
        r = np.linalg.norm(B[:, c] - A @ X[:, c])
        assert abs(np.linalg.norm(QtB[n:, c]) - r) <= 1e-12 * np.linalg.norm(B[:, c]) * m


def test_dgels_matches_oracle(h, orc):
    rng = np.random.default_rng(7)
    for (m, n, nrhs) in [(600, 200, 1), (1000, 1000, 2), (4000, 130, 3)]:
        note(f"dgels {m}x{n} nrhs={nrhs}")
        A = rng.standard_normal((m, n))
        B = rng.standard_normal((m, nrhs))
        dA, dB = M.colmajor(A), M.colmajor(B)
        X = host(M.gels(h, dA, dB))
        for c in range(nrhs):
            xo = orc.gels(A, B[:, c])
            np.testing.assert_allclose(X[:, c], xo, rtol=0, atol=1e-9 * max(1, np.abs(xo).max()))


# ---------------------------------------------------------------------------
# special cases (closed forms) and the host-level drop-in
# ---------------------------------------------------------------------------

@pytest.mark.parametrize("n", [1, 5, 64, 200])
def test_identity_gives_minus_identity_exactly(h, n):
    dA, tau, T = gpu_factor(h, np.eye(n))
    dQ = M.orgqr(h, dA, T, n, nq=n)
    h.synchronize()
    assert np.array_equal(np.triu(host(dA)), -np.eye(n))
    assert np.array_equal(host(dQ), -np.eye(n))
    assert np.all(tau.cpu().numpy() == 2.0)


def test_zero_and_dependent_columns(h, orc):
    rng = np.random.default_rng(8)
    for (m, n, zc) in [(8, 5, 0), (400, 250, 0), (400, 250, 130), (300, 300, 299)]:
        A = rng.standard_normal((m, n))
        A[:, zc] = 0.0
        dA, tau, T = gpu_factor(h, A)
        tau = tau.cpu().numpy()
        if zc == 0:
            assert tau[0] == 0.0
        VR = host(dA)
        dQ = M.orgqr(h, dA, T, min(m, n), nq=m)
        Q = host(dQ)
        assert np.linalg.norm(Q @ np.triu(VR) - A) < 1e-12 * m
        assert np.linalg.norm(Q.T @ Q - np.eye(m)) < 1e-12 * m


def haar_np(rng, rows, cols):
    Q, R = np.linalg.qr(rng.standard_normal((rows, cols)))
    return Q * np.sign(np.diag(R))[None, :]


@pytest.mark.parametrize("m,n,smin", [(1500, 1000, 1e-8), (3000, 700, 1e-12), (1200, 1200, 1e-14)])
def test_graded_matrices_match_oracle(h, orc, m, n, smin):
    """Tall panels (CholeskyQR-based) on ill-conditioned inputs: same R, V,
    tau as the oracle to backward-error level, orthogonal Q."""
    note(f"graded {m}x{n} smin={smin}")
    rng = np.random.default_rng(int(m + n))
    A = (haar_np(rng, m, n) * np.logspace(0, np.log10(smin), n)) @ haar_np(rng, n, n).T
    dA, tau, T = gpu_factor(h, A)
    dQ = M.orgqr(h, dA, T, n, nq=n)
    VR, Q = host(dA), host(dQ)
    R = np.triu(VR)[:n]
    VRo, _, tau_o = orc.geqrf_blocked(A)
    nA = np.linalg.norm(A)
    assert np.linalg.norm(Q @ R - A) / nA < 8 * m * EPS
    assert np.linalg.norm(Q.T @ Q - np.eye(n)) < 8 * m * EPS
    # R agrees entrywise with the oracle relative to ||A|| (same sign convention)
    assert np.abs(R - np.triu(VRo)[:n]).max() / nA < 1e-12
    # reflectors agree where the columns are well determined (leading block)
    lead = n // 4
    np.testing.assert_allclose(tau[:lead].cpu().numpy(), tau_o[:lead], rtol=0, atol=1e-10)


def test_tall_panel_fallback_on_dependent_columns(h, orc):
    """Zero and duplicated columns inside tall panels: the fast panel hands
    over to the column-by-column kernel, which keeps tau = 0 for a column that
    is exactly zero, as the reference does."""
    rng = np.random.default_rng(11)
    m, n = 2000, 300
    A = rng.standard_normal((m, n))
    A[:, 5] = 0.0                 # zero column inside the first panel
    A[:, 70] = A[:, 66]           # duplicate inside the second panel
    A[:, 200] = 2.0 * A[:, 3]     # dependent on an earlier panel
    dA, tau, T = gpu_factor(h, A)
    tau = tau.cpu().numpy()
    assert tau[5] == 0.0
    VR = host(dA)
    assert np.all(VR[5:, 5] == 0.0)
    dQ = M.orgqr(h, dA, T, n, nq=m)
    Q = host(dQ)
    assert np.linalg.norm(Q @ np.triu(VR) - A) / np.linalg.norm(A) < 1e-14 * m
    assert np.linalg.norm(Q.T @ Q - np.eye(m)) < 1e-13 * m
    VRo, _, tau_o = orc.geqrf_blocked(A)
    assert tau_o[5] == 0.0
    assert abs(VR[70, 70]) < 1e-12 * np.linalg.norm(A) and abs(VR[200, 200]) < 1e-12 * np.linalg.norm(A)


@pytest.mark.parametrize("nb", [16, 32, 48, 128, 192, 256])
def test_block_sizes_match_oracle(torch_mod, orc, nb):
    """Every block width mqr_set_block_size takes: below 64 the panels are nb
    wide; above, the panels are 64 wide and each nb-column block is applied
    to the far columns as one block reflector, whose T (the larft recurrence
    over all nb reflectors) is what the T blocks hold.  The same R and Q as
    the oracle, and each stored T block equal to larft of its reflectors (Q
    formation then applies the stored blocks)."""
    note(f"block size {nb}")
    with M.Handle(0, nb=nb) as hn:
        for (m, n) in [(300, 200), (700, 150), (900, 700)]:
            rng = np.random.default_rng(m + n + nb)
            A = rng.standard_normal((m, n))
            dA, tau, T = gpu_factor(hn, A)
            dQ = M.orgqr(hn, dA, T, n, nq=n)
            hn.synchronize()
            VR, Q, T, tau = host(dA), host(dQ), host(T), tau.cpu().numpy()
            Qo, Ro, _ = orc.qr(A)
            nA = np.linalg.norm(A)
            R = np.triu(VR)[:n]
            np.testing.assert_allclose(R, Ro[:n], rtol=0, atol=1e-12 * nA)
            np.testing.assert_allclose(Q, Qo[:, :n], rtol=0, atol=1e-12 * m)
            for j in range(0, n, nb):
                b = min(nb, n - j)
                Tref = orc.larft(unit_lower(VR[j:, j:j + b], b), tau[j:j + b])
                np.testing.assert_allclose(np.triu(T[:b, j:j + b]), Tref, rtol=0,
                                           atol=64 * m * EPS * max(1.0, np.abs(Tref).max()))


def test_short_panel_fallback_on_dependent_columns(h, orc):
    """The same hand-over in the last panels of a square matrix (64 to 128
    rows left): a zero column stays exactly zero under the earlier block
    reflectors and keeps tau = 0; a duplicated column leaves a zero pivot.
    R agrees with the oracle in its rows above that pivot; below it the
    reflectors follow rounding noise, and any rotation of those rows is an
    equally valid factorisation."""
    rng = np.random.default_rng(12)
    m = n = 320
    A = rng.standard_normal((m, n))
    A[:, 200] = 0.0               # panel at column 192: 128 rows left
    A[:, 270] = A[:, 260]         # last panel: 64 rows left
    dA, tau, T = gpu_factor(h, A)
    tau = tau.cpu().numpy()
    assert tau[200] == 0.0
    VR = host(dA)
    assert np.all(VR[200:, 200] == 0.0)
    dQ = M.orgqr(h, dA, T, n, nq=m)
    Q = host(dQ)
    nA = np.linalg.norm(A)
    R = np.triu(VR)
    assert np.linalg.norm(Q @ R - A) / nA < 1e-14 * m
    assert np.linalg.norm(Q.T @ Q - np.eye(m)) < 1e-13 * m
    VRo, _, tau_o = orc.geqrf_blocked(A)
    assert tau_o[200] == 0.0
    assert abs(VR[270, 270]) < 1e-12 * nA
    np.testing.assert_allclose(R[:270], np.triu(VRo)[:270], rtol=0, atol=1e-12 * nA)


def test_host_qr_api_matches_oracle(orc):
    rng = np.random.default_rng(9)
    for (m, n) in [(7, 4), (300, 200), (257, 300)]:
        A = np.asfortranarray(rng.standard_normal((m, n)))
        Q = np.zeros((m, m), order="F")
        R = np.zeros((m, n), order="F")
        M.QR(A, Q, R, M.QR_Householder)
        Qo, Ro, _ = orc.qr(A)
        assert np.all(np.tril(R, -1) == 0.0)
        np.testing.assert_allclose(R, np.triu(Ro), rtol=0, atol=1e-12 * max(m, n))
        np.testing.assert_allclose(Q, Qo, rtol=0, atol=1e-12 * m)


def test_argument_errors_are_lapack_style(h):
    L = M.lib()
    assert L.mqr_dgeqrf(h.h, -1, 4, None, 4, None, None, 64) == -2
    assert L.mqr_dgeqrf(h.h, 4, -1, None, 4, None, None, 64) == -3
    assert L.mqr_dgeqrf(h.h, 4, 4, None, 4, None, None, 64) == -4
    import torch
    A = torch.zeros(16, device="cuda", dtype=torch.float64)
    assert L.mqr_dgeqrf(h.h, 4, 4, M._ptr(A), 3, None, None, 64) == -5
    assert L.mqr_dgeqrf(h.h, 4, 4, M._ptr(A), 4, None, None, 64) == -6
    assert L.mqr_set_block_size(h.h, 24) == -2
    assert L.mqr_dorgqr(h.h, 4, 5, 4, M._ptr(A), 4, M._ptr(A), 64, M._ptr(A), 4) == -3
    assert L.mqr_dgels(h.h, 3, 4, 1, M._ptr(A), 3, M._ptr(A), 3) == -3


@pytest.mark.parametrize("bad", [np.nan, np.inf])
def test_nonfinite_entry_propagates_like_the_reference(h, orc, bad):
    """A NaN or Inf in column 70 (the reference checks nothing and lets it run
    through its arithmetic): the factorisation of the first 70 columns is
    untouched, R and tau there equal the oracle's, and the value reaches R
    from row 70 on in the later columns.  The fast panel holding it hands over
    to the column-by-column factorisation."""
    note(f"non-finite entry {bad}")
    rng = np.random.default_rng(13)
    m, n = 300, 200
    A = rng.standard_normal((m, n))
    A[10, 70] = bad
    dA, tau, T = gpu_factor(h, A)
    VR, tau = host(dA), tau.cpu().numpy()
    R = np.triu(VR)[:n]
    assert np.all(np.isfinite(R[:, :70])) and np.all(np.isfinite(tau[:70]))
    assert not np.all(np.isfinite(R[70:, 70:]))
    _, Ro, _ = orc.qr(A[:, :70])
    np.testing.assert_allclose(R[:70, :70], np.triu(Ro)[:70], rtol=0,
                               atol=1e-12 * np.linalg.norm(A[:, :70]))


def test_t_blocks_rejected_after_block_size_change(torch_mod):
    """T is laid out for the block size of its factorisation: forming or
    applying Q after mqr_set_block_size changed it is refused (LAPACK-style
    position of T), and accepted again once the block size is restored."""
    rng = np.random.default_rng(14)
    A = rng.standard_normal((200, 150))
    with M.Handle(0, nb=64) as hn:
        dA, tau, T = gpu_factor(hn, A)
        L = M.lib()
        Q = M.empty_colmajor(200, 150)
        M.check(L.mqr_set_block_size(hn.h, 32), "mqr_set_block_size")
        assert L.mqr_dorgqr(hn.h, 200, 150, 150, M._ptr(dA), 200, M._ptr(T), M.ld(T),
                            M._ptr(Q), 200) == -7
        assert L.mqr_dormqr(hn.h, b"T", 200, 150, 150, M._ptr(dA), 200, M._ptr(T), M.ld(T),
                            M._ptr(Q), 200) == -8
        M.check(L.mqr_set_block_size(hn.h, 64), "mqr_set_block_size")
        Q = M.orgqr(hn, dA, T, 150, nq=150)
        hn.synchronize()
        assert np.linalg.norm(host(Q) @ np.triu(host(dA))[:150] - A) < 1e-12 * np.linalg.norm(A)


def test_host_gemm_behind_operator_star(orc):
    """mqr_dgemm_host / mqr_sgemm_host (the product MAT<T>::operator* runs):
    exact on integer-valued operands against the oracle's packed GEMM
    (Mul/Mul.h:239-263), within the FP64 / FP32 error bound on Gaussian
    ones; k = 0 gives zeros."""
    rng = np.random.default_rng(15)
    for (m, n, k) in [(1, 1, 1), (5, 3, 4), (130, 70, 33), (300, 257, 129)]:
        A = rng.integers(-8, 9, (m, k)).astype(np.float64)
        B = rng.integers(-8, 9, (k, n)).astype(np.float64)
        exact = orc.mulAB_packed_n(A, B, np.zeros((m, n)))
        assert np.array_equal(M.gemm_host(A, B), -exact)
        assert np.array_equal(M.gemm_host(A.astype(np.float32), B.astype(np.float32)),
                              (-exact).astype(np.float32))
    A = rng.standard_normal((400, 200))
    B = rng.standard_normal((200, 90))
    ref = orc.gemm(A, B)
    np.testing.assert_allclose(M.gemm_host(A, B), ref, rtol=0, atol=200 * 4 * EPS * 20)
    np.testing.assert_allclose(M.gemm_host(A.astype(np.float32), B.astype(np.float32)), ref,
                               rtol=0, atol=200 * 4 * 6e-8 * 20)
    assert np.array_equal(M.gemm_host(np.zeros((4, 0)), np.zeros((0, 3))), np.zeros((4, 3)))
