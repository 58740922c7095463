"""Pins the CPU oracle (oracle/mqr_oracle.c) before anything is checked
against it.

The reference ships no test vectors (no test sources, logs without inputs), so
the oracle is pinned by: LAPACK vectors generated with scipy
(tests/golden/lapack_qr.npz), brute-force products of explicit reflectors on
small inputs, closed forms (identity, 1x1, zero columns), exact integer
arithmetic for the packed GEMM, and invariants at the 1024 x 1024 size.  The
last tests plant mistakes in oracle outputs and check that the checks catch
them.
"""
import numpy as np
import pytest
import scipy.linalg as sla

from oracle.oracle import Oracle, ls_backward_error, qr_adjoint

VARIANTS = ["avx2", "avx512"]
EPS = np.finfo(np.float64).eps


@pytest.fixture(scope="module", params=VARIANTS)
def orc(request):
    return Oracle(request.param)


def _shapes(golden):
    out = []
    for key in golden.files:
        if key.startswith("A_"):
            m, n = key[2:].split("x")
            out.append((int(m), int(n)))
    return sorted(out)


def _householders(VR, tau):
    """Explicit reflectors H_r = I - tau_r v_r v_r^T from packed storage."""
    m = VR.shape[0]
    Hs = []
    for r in range(len(tau)):
        v = np.zeros(m)
        v[r] = 1.0
        v[r + 1:] = VR[r + 1:, r]
        Hs.append(np.eye(m) - tau[r] * np.outer(v, v))
    return Hs


def _unit_lower(VR, b):
    V = np.tril(VR[:, :b], -1)
    V[np.arange(b), np.arange(b)] = 1.0
    return V


def _t_residual(V, tau, T):
    m, b = V.shape
    H = np.eye(m)
    for Hr in _householders(np.column_stack([V, np.zeros((m, 0))]), tau):
        H = H @ Hr
    return np.abs(np.eye(m) - V @ T @ V.T - H).max()


# ---------------------------------------------------------------------------
# LAPACK vectors
# ---------------------------------------------------------------------------

def test_blocked_factor_matches_lapack(orc, golden):
    for (m, n) in _shapes(golden):
        key = f"{m}x{n}"
        A = golden[f"A_{key}"]
        VRl, taul = golden[f"VR_{key}"], golden[f"tau_{key}"]
        VR, _, tau = orc.geqrf_blocked(A)
        k = min(m, n)
        tol = 64 * max(m, n) * EPS * max(1.0, np.abs(A).max())
        kk = k - 1 if m <= n else k  # last reflector of length 1: reference keeps tau = 2
        np.testing.assert_allclose(tau[:kk], taul[:kk], rtol=0, atol=tol, err_msg=key)
        np.testing.assert_allclose(np.triu(VR)[:kk], np.triu(VRl)[:kk], rtol=0, atol=tol, err_msg=key)
        np.testing.assert_allclose(np.tril(VR, -1), np.tril(VRl, -1), rtol=0, atol=tol, err_msg=key)
        if m <= n:
            # reference reflects a length-1 column anyway: tau = 2, row negated
            assert tau[k - 1] == 2.0 or (VRl[k - 1, k - 1] == 0.0 and tau[k - 1] == 0.0)
            if tau[k - 1] == 2.0:
                np.testing.assert_allclose(VR[k - 1, k - 1:], -VRl[k - 1, k - 1:], rtol=0, atol=tol)


def test_larft_matches_lapack_T(orc, golden):
    for (m, n) in _shapes(golden):
        key = f"{m}x{n}"
        k = min(m, n)
        V = _unit_lower(golden[f"VR_{key}"], k)
        T = orc.larft(V, golden[f"tau_{key}"])
        Tl = golden[f"T_{key}"]
        np.testing.assert_allclose(np.triu(T), np.triu(Tl), rtol=0,
                                   atol=64 * m * EPS * max(1, np.abs(Tl).max()), err_msg=key)


def test_full_qr_matches_lapack_q(orc, golden):
    for (m, n) in _shapes(golden):
        key = f"{m}x{n}"
        A = golden[f"A_{key}"]
        Q, R, _ = orc.qr(A)
        Ql = golden[f"Q_{key}"]
        k = min(m, n)
        Qk = Q[:, :k].copy()
        if m <= n and m > 0:
            Qk[:, k - 1] *= -1.0  # H_k = diag(I, -1) in the reference
        np.testing.assert_allclose(Qk, Ql, rtol=0, atol=64 * m * EPS, err_msg=key)
        np.testing.assert_allclose(Q @ R, A, rtol=0, atol=64 * max(m, n) * EPS * np.abs(A).max())


# ---------------------------------------------------------------------------
# brute force on small inputs
# ---------------------------------------------------------------------------

@pytest.mark.parametrize("h,b", [(1, 1), (4, 4), (9, 3), (20, 7), (33, 16)])
def test_panel_is_product_of_reflectors(orc, h, b):
    rng = np.random.default_rng(h * 100 + b)
    P = rng.standard_normal((h, b))
    VR, tau = orc.panel(P)
    Hs = _householders(VR, tau)
    H = np.eye(h)
    for Hr in Hs:
        H = H @ Hr
        np.testing.assert_allclose(Hr @ Hr.T, np.eye(h), atol=1e-13)  # each H orthogonal
    Rfull = np.triu(VR)
    np.testing.assert_allclose(H.T @ P, Rfull, atol=1e-13 * max(1, np.abs(P).max()))
    # sign convention: R_rr = -sign(alpha_r) * norm, so R_rr * alpha_r <= 0
    for r in range(min(h, b)):
        Pr = P.copy()
        for Hr in Hs[:r]:
            Pr = Hr @ Pr
        assert Rfull[r, r] * Pr[r, r] <= 0.0


@pytest.mark.parametrize("h,b", [(6, 3), (25, 8), (40, 17)])
def test_larft_brute_force(orc, h, b):
    rng = np.random.default_rng(7 * h + b)
    VR, tau = orc.panel(rng.standard_normal((h, b)))
    V = _unit_lower(VR, b)
    T = orc.larft(V, tau)
    assert np.allclose(np.tril(T, -1), 0)
    assert _t_residual(V, tau, T) < 1e-13


@pytest.mark.parametrize("h,b,nc", [(10, 4, 3), (37, 9, 11), (150, 20, 160), (300, 96, 7)])
def test_larfb_matches_sequential_reflectors(orc, h, b, nc):
    rng = np.random.default_rng(h + 31 * b + nc)
    b = min(b, orc.BS)
    VR, tau = orc.panel(rng.standard_normal((h, b)))
    V = _unit_lower(VR, b)
    T = orc.larft(V, tau)
    Cm = rng.standard_normal((h, nc))
    Hs = _householders(VR, tau)
    ref_t = Cm.copy()
    for Hr in Hs:
        ref_t = Hr @ ref_t
    ref_n = Cm.copy()
    for Hr in reversed(Hs):
        ref_n = Hr @ ref_n
    np.testing.assert_allclose(orc.larfb(V, T, Cm, transposed=True), ref_t, atol=1e-12)
    np.testing.assert_allclose(orc.larfb(V, T, Cm, transposed=False), ref_n, atol=1e-12)


def test_packed_gemm_exact_on_integers(orc):
    rng = np.random.default_rng(3)
    for (m, n, k) in [(1, 1, 1), (7, 5, 3), (orc.MR, orc.NR, 9), (37, 29, 41), (200, 150, 90)]:
        A = rng.integers(-8, 9, (m, k)).astype(np.float64)
        B = rng.integers(-8, 9, (k, n)).astype(np.float64)
        Cm = rng.integers(-8, 9, (m, n)).astype(np.float64)
        exact = Cm.astype(np.int64) - A.astype(np.int64) @ B.astype(np.int64)
        assert np.array_equal(orc.mulAB_packed_n(A, B, Cm), exact.astype(np.float64))
    for (M, N, K) in [(5, 3, 2), (600, 170, 300)]:
        A = rng.integers(-8, 9, (M, K)).astype(np.float64)
        B = rng.integers(-8, 9, (K, N)).astype(np.float64)
        assert np.array_equal(orc.gemm(A, B), (A.astype(np.int64) @ B.astype(np.int64)).astype(np.float64))


def test_pack_layouts_round_trip(orc):
    rng = np.random.default_rng(4)
    A = rng.standard_normal((2 * orc.MR + 3, 5))
    Ap = orc.pack_A(A)
    mc, kc = A.shape
    for ip in range((mc + orc.MR - 1) // orc.MR):
        panel = Ap[ip * kc * orc.MR:(ip + 1) * kc * orc.MR].reshape(kc, orc.MR).T
        rows = A[ip * orc.MR:(ip + 1) * orc.MR]
        np.testing.assert_array_equal(panel[:rows.shape[0]], rows)
        assert np.all(panel[rows.shape[0]:] == 0)
    B = rng.standard_normal((6, orc.NR + 2))
    Bp = orc.pack_B(B)
    kc, nc = B.shape
    for jp in range((nc + orc.NR - 1) // orc.NR):
        panel = Bp[jp * kc * orc.NR:(jp + 1) * kc * orc.NR].reshape(kc, orc.NR)
        cols = B[:, jp * orc.NR:(jp + 1) * orc.NR]
        np.testing.assert_array_equal(panel[:, :cols.shape[1]], cols)


# ---------------------------------------------------------------------------
# closed forms and special cases
# ---------------------------------------------------------------------------

@pytest.mark.parametrize("n", [1, 5, 150, 300])
def test_identity_gives_minus_identity(orc, n):
    Q, R, blocked = orc.qr(np.eye(n))
    if blocked:  # tau = 2, v = 0: exact
        np.testing.assert_array_equal(R, -np.eye(n))
        np.testing.assert_array_equal(Q, -np.eye(n))
    else:  # H = I - w w^T with w = sqrt(2) e_r: -1 up to the rounding of sqrt(2)^2
        np.testing.assert_allclose(R, -np.eye(n), rtol=0, atol=4 * EPS)
        np.testing.assert_allclose(Q, -np.eye(n), rtol=0, atol=4 * EPS)


def test_one_by_one(orc):
    for a in [3.5, -2.0]:
        Q, R, _ = orc.qr(np.array([[a]]))  # unblocked branch: w = sqrt(2)
        assert abs(R[0, 0] + a) <= 4 * EPS * abs(a) and abs(Q[0, 0] + 1.0) <= 4 * EPS
        VR, _, tau = orc.geqrf_blocked(np.array([[a]]))  # blocked form: exact
        assert VR[0, 0] == -a and tau[0] == 2.0
    Q, R, _ = orc.qr(np.zeros((1, 1)))
    assert R[0, 0] == 0.0 and Q[0, 0] == 1.0


def test_zero_column_gives_zero_tau(orc):
    rng = np.random.default_rng(5)
    for (m, n, zc) in [(8, 5, 0), (8, 5, 2), (400, 250, 0), (400, 250, 130)]:
        A = rng.standard_normal((m, n))
        A[:, zc] = 0.0
        if zc > 0:
            A[:, zc] = A[:, :zc] @ rng.standard_normal(zc)  # dependent: zero after projection
        VR, _, tau = orc.geqrf_blocked(A)
        if zc == 0:
            assert tau[0] == 0.0 and np.all(VR[:, 0] == 0.0)
        assert abs(VR[zc, zc]) < 1e-12 * np.abs(A).max() * m
        Q, R, _ = orc.qr(A)
        assert np.allclose(Q @ R, A, atol=1e-12 * m)
        assert np.allclose(Q.T @ Q, np.eye(m), atol=1e-12 * m)


def test_wide_matrices(orc):
    rng = np.random.default_rng(6)
    for (m, n) in [(3, 7), (50, 90), (250, 400)]:
        A = rng.standard_normal((m, n))
        Q, R, _ = orc.qr(A)
        assert np.allclose(Q @ R, A, atol=1e-12 * n)
        assert np.abs(np.tril(R, -1)).max() < 1e-12 * n


def test_unblocked_and_blocked_agree(orc):
    rng = np.random.default_rng(8)
    A = rng.standard_normal((260, 230))
    Qu, Ru = orc.qr_unblocked(A)
    VR, T_all, _ = orc.geqrf_blocked(A)
    Qb = orc.orgqr_full(VR, T_all, 230)
    np.testing.assert_allclose(np.triu(Ru), np.triu(VR), atol=1e-11)
    np.testing.assert_allclose(Qu, Qb, atol=1e-12)


def test_variants_agree():
    a, b = Oracle("avx2"), Oracle("avx512")
    rng = np.random.default_rng(9)
    A = rng.standard_normal((500, 300))
    Qa, Ra, _ = a.qr(A)
    Qb, Rb, _ = b.qr(A)
    np.testing.assert_allclose(Ra, Rb, atol=1e-11)
    np.testing.assert_allclose(Qa, Qb, atol=1e-12)


@pytest.mark.parametrize("isa", ["avx2", "avx512"])
def test_simd_builds_are_bit_identical(isa):
    """The register-accumulator micro-kernel (the timed CPU baseline) rounds
    exactly like the scalar restatement: same fused multiply-adds in the same
    k order per element.  Edge tiles (300 = 18 MR + 12 rows for 16-row tiles,
    211 columns) take the scalar path inside the same call."""
    if isa == "avx512":
        from oracle.oracle import cpu_has_avx512
        if not cpu_has_avx512():
            pytest.skip("host has no AVX-512")
    a, b = Oracle(isa), Oracle(isa + "_simd")
    rng = np.random.default_rng(21)
    A = rng.standard_normal((300, 170))
    B = rng.standard_normal((170, 211))
    assert np.array_equal(a.gemm(A, B), b.gemm(A, B))
    M = np.asfortranarray(rng.standard_normal((700, 300)))
    y = rng.standard_normal(700)
    assert np.array_equal(a.gels(M, y), b.gels(M, y))
    Qa, Ra, _ = a.qr(M[:, :250])
    Qb, Rb, _ = b.qr(M[:, :250])
    assert np.array_equal(Ra, Rb) and np.array_equal(Qa, Qb)


def test_c1_invariants_and_lapack_r(golden):
    """The CPU configuration: 1024 x 1024 iid N(0,1), blocked branch."""
    o = Oracle("avx2")
    rng = np.random.default_rng(1024)
    A = rng.standard_normal((1024, 1024))
    Q, R, blocked = o.qr(A)
    assert blocked
    assert np.linalg.norm(Q @ R - A) / np.linalg.norm(A) < 1e-14
    assert np.linalg.norm(Q.T @ Q - np.eye(1024)) < 1e-12
    assert np.all(np.tril(R, -1) == 0.0)
    Rl = sla.qr(A, mode="r")[0]
    s = np.sign(np.diag(R)) * np.sign(np.diag(Rl))
    np.testing.assert_allclose(s[:, None] * R, Rl, atol=1e-11)
    assert np.all(s[:-1] == 1.0) and s[-1] == -1.0  # only the length-1 reflector differs


# ---------------------------------------------------------------------------
# least squares, backward error and adjoint helpers (no reference counterpart)
# ---------------------------------------------------------------------------

def test_gels_matches_lstsq(orc):
    rng = np.random.default_rng(10)
    A = rng.standard_normal((300, 120))
    b = rng.standard_normal(300)
    x = orc.gels(A, b)
    xl = np.linalg.lstsq(A, b, rcond=None)[0]
    np.testing.assert_allclose(x, xl, atol=1e-12)


def test_backward_error_estimate_against_exact_formula():
    """Karlson-Walden estimate vs Walden-Karlson-Sun's exact normwise backward
    error eta = min(phi, sigma_min([A, phi (I - r r^+)])), phi = ||r||/||x||."""
    rng = np.random.default_rng(11)
    A = rng.standard_normal((12, 5))
    b = rng.standard_normal(12)
    x0 = np.linalg.lstsq(A, b, rcond=None)[0]
    for scale in [1e-3, 1e-6, 1e-9]:
        x = x0 + scale * rng.standard_normal(5)
        r = b - A @ x
        phi = np.linalg.norm(r) / np.linalg.norm(x)
        M = np.hstack([A, phi * (np.eye(12) - np.outer(r, r) / (r @ r))])
        eta = min(phi, np.linalg.svd(M, compute_uv=False).min())
        kw = ls_backward_error(A, b, x)["kw"] * np.linalg.norm(A)
        assert 0.5 < kw / eta < 2.0, (scale, kw, eta)


def test_adjoint_formula_against_finite_differences():
    rng = np.random.default_rng(12)
    m, n = 9, 4
    A = rng.standard_normal((m, n))
    Wq = rng.standard_normal((m, n))
    Wr = rng.standard_normal((n, n))

    def loss(X):
        Q, R = np.linalg.qr(X)
        return np.sum(Wq * Q) + np.sum(Wr * np.triu(R))

    Q, R = np.linalg.qr(A)
    dA = qr_adjoint(Q, R, Wq, np.triu(Wr))
    h = 1e-6
    fd = np.zeros_like(A)
    for i in range(m):
        for j in range(n):
            E = np.zeros_like(A)
            E[i, j] = h
            fd[i, j] = (loss(A + E) - loss(A - E)) / (2 * h)
    np.testing.assert_allclose(dA, fd, atol=1e-7)


# ---------------------------------------------------------------------------
# planted mistakes: the checks above must reject them
# ---------------------------------------------------------------------------

def test_planted_mistakes_are_caught():
    o = Oracle("avx2")
    rng = np.random.default_rng(13)
    VR, tau = o.panel(rng.standard_normal((30, 6)))
    V = _unit_lower(VR, 6)
    T = o.larft(V, tau)
    assert _t_residual(V, tau, T) < 1e-13
    bad_tau = tau.copy()
    bad_tau[2] = -bad_tau[2]
    assert _t_residual(V, tau, o.larft(V, bad_tau)) > 1e-3      # wrong sign of a tau
    assert _t_residual(V, tau, T.T) > 1e-3                       # transposed T
    T_drop = T.copy()
    T_drop[0, 3] = 0.0
    assert _t_residual(V, tau, T_drop) > 1e-6                    # dropped coupling term
    Cm = rng.standard_normal((30, 4))
    good = o.larfb(V, T, Cm, transposed=True)
    assert not np.allclose(good, o.larfb(V, T, Cm, transposed=False), atol=1e-6)  # Q vs Q^T


def test_restatement_under_sanitizers():
    """The C restatement built with AddressSanitizer and UndefinedBehaviorSanitizer
    (`make -C oracle sanitize`) runs its edge-case driver (oracle/sanity_main.c:
    1x1, wide, both branches of the QR call, zero columns, least squares, the
    packed GEMM) with no out-of-bounds access or undefined operation, and every
    reconstruction check in it passes."""
    import pathlib
    import shutil
    import subprocess
    from oracle.oracle import cpu_has_avx512
    if shutil.which("make") is None or (shutil.which("cc") is None and shutil.which("gcc") is None):
        pytest.skip("no C toolchain")
    root = pathlib.Path(__file__).resolve().parents[1] / "oracle"
    r = subprocess.run(["make", "-s", "-C", str(root), "sanitize"], capture_output=True, text=True,
                       timeout=600)
    assert r.returncode == 0, r.stdout + r.stderr
    for v in ["avx2"] + (["avx512"] if cpu_has_avx512() else []):
        r = subprocess.run([str(root / "lib" / f"orc_sanity_{v}")], capture_output=True, text=True,
                           timeout=600)
        assert r.returncode == 0, r.stdout + r.stderr
        assert "all checks passed" in r.stdout


# ---------------------------------------------------------------------------
# the other QR_type values: Givens (QR/__Givens.h:45-63) and modified
# Gram-Schmidt (QR/__GramSchmidt.h:2-18)
# ---------------------------------------------------------------------------

def _row_signs(R, Rl, k):
    d = np.sign(np.diag(R)[:k]) * np.sign(np.diag(Rl)[:k])
    d[d == 0] = 1.0
    return d


@pytest.mark.parametrize("m,n", [(2, 1), (7, 4), (40, 40), (90, 33), (25, 60)])
def test_givens_invariants_and_lapack_r(orc, m, n):
    rng = np.random.default_rng(m * 31 + n)
    A = rng.standard_normal((m, n))
    Q, R = orc.qr_givens(A)
    nA = np.linalg.norm(A)
    k = min(m, n)
    assert np.linalg.norm(Q @ R - A) / nA < 16 * max(m, n) * EPS
    assert np.linalg.norm(Q.T @ Q - np.eye(m)) < 16 * m * EPS
    # below the diagonal only the rotations' rounding residue remains
    assert np.abs(np.tril(R, -1)).max() < 16 * max(m, n) * EPS * nA
    # R is unique up to row signs: equal to LAPACK's after fixing them
    Rl = np.triu(sla.qr(A, mode="r")[0])
    d = _row_signs(R, Rl, k)
    np.testing.assert_allclose(d[:, None] * np.triu(R)[:k], Rl[:k], rtol=0,
                               atol=64 * max(m, n) * EPS * nA)
    # each column's sweep keeps the sign of its pivot: R(0,0) = sign(a00) ||a_0||
    assert np.sign(R[0, 0]) == np.sign(A[0, 0])
    assert abs(abs(R[0, 0]) - np.linalg.norm(A[:, 0])) < 8 * m * EPS * nA


def test_givens_zero_pivots(orc):
    """The source divides by a zero pivot (NaN spreads); the GPU path's
    documented fix gives exact results on triangular inputs."""
    A = np.triu(np.random.default_rng(5).standard_normal((6, 4)))
    Q, R = orc.qr_givens(A)
    assert np.isnan(R).any() or np.isnan(Q).any()
    Q, R = orc.qr_givens(A, fix_zero_pivots=True)
    assert np.array_equal(R, A) and np.array_equal(Q, np.eye(6))
    # a zero pivot over a nonzero entry: rows swapped with the sign chosen so
    # that the new pivot is |b|
    B = np.array([[0.0, 1.0], [-3.0, 2.0]])
    Q, R = orc.qr_givens(B, fix_zero_pivots=True)
    assert R[0, 0] == 3.0 and R[1, 0] == 0.0
    np.testing.assert_allclose(Q @ R, B, atol=1e-15)
    np.testing.assert_allclose(Q.T @ Q, np.eye(2), atol=1e-15)


@pytest.mark.parametrize("m,n", [(1, 1), (7, 4), (40, 40), (300, 70)])
def test_gramschmidt_invariants_and_positive_r(orc, m, n):
    rng = np.random.default_rng(m * 17 + n)
    A = rng.standard_normal((m, n))
    Q, R = orc.qr_gramschmidt(A)
    nA = np.linalg.norm(A)
    assert Q.shape == (m, n) and R.shape == (m, n)
    assert np.all(R[n:] == 0.0) and np.all(np.tril(R, -1) == 0.0)
    assert np.all(np.diag(R) > 0.0)
    assert np.linalg.norm(Q @ R[:n] - A) / nA < 16 * m * EPS
    assert np.linalg.norm(Q.T @ Q - np.eye(n)) < 64 * m * EPS * np.linalg.cond(A)
    # the thin QR with a positive diagonal is unique: LAPACK's, signs fixed
    Rl = np.triu(sla.qr(A, mode="economic")[1])
    Rl = np.sign(np.diag(Rl))[:, None] * Rl
    np.testing.assert_allclose(R[:n], Rl, rtol=0, atol=64 * m * EPS * nA * np.linalg.cond(A))


def test_gramschmidt_is_the_modified_variant(orc):
    """Lauchli's matrix with eps^2 below the unit roundoff: classical
    Gram-Schmidt loses orthogonality completely (q2^T q3 = 1/2), the
    modified form of the source keeps it at O(eps)."""
    e = 1e-9
    A = np.array([[1.0, 1.0, 1.0], [e, 0, 0], [0, e, 0], [0, 0, e]])
    Q, _ = orc.qr_gramschmidt(A)
    assert abs(Q[:, 1] @ Q[:, 2]) < 1e-6
    # the planted classical variant (every projection against the original
    # column) is caught by the same check
    Qc = np.zeros_like(A)
    for j in range(3):
        v = A[:, j] - Qc[:, :j] @ (Qc[:, :j].T @ A[:, j])
        Qc[:, j] = v / np.linalg.norm(v)
    assert abs(Qc[:, 1] @ Qc[:, 2]) > 0.4


def test_planted_mistakes_in_givens_and_gs_are_caught():
    o = Oracle("avx2")
    A = np.random.default_rng(21).standard_normal((12, 7))
    Q, R = o.qr_givens(A)
    assert np.linalg.norm(Q @ R - A) < 1e-13
    assert np.linalg.norm(Q.T @ R - A) > 1e-2          # Q transposed
    Rb = R.copy()
    Rb[3] = -Rb[3]
    assert np.linalg.norm(Q @ Rb - A) > 1e-2           # one row sign flipped
    Qg, Rg = o.qr_gramschmidt(A)
    Rg_bad = Rg.copy()
    Rg_bad[1, 4] = 0.0                                  # dropped projection coefficient
    assert np.linalg.norm(Qg @ Rg_bad[:7] - A) > 1e-3


# ---------------------------------------------------------------------------
# LU / PLU / Solve_direct (LU/__LU.h, LU/__PLU.h, LU/__Solve_direct.h)
# ---------------------------------------------------------------------------

@pytest.mark.parametrize("m,n", [(1, 1), (5, 5), (40, 40), (90, 33), (33, 90), (300, 300)])
def test_plu_matches_lapack_getrf(orc, m, n):
    """Partial pivoting picks the first row of largest magnitude, as LAPACK's
    dgetrf (idamax) does: the same permutation, and L, U equal to rounding."""
    rng = np.random.default_rng(m * 5 + n)
    A = rng.standard_normal((m, n))
    F, pos, done = orc.lu(A)
    assert done == min(m, n)
    lu, piv = sla.lu_factor(A) if m == n else (None, None)
    # LAPACK's sequential swaps -> the permutation vector
    import scipy.linalg.lapack as la
    Fl, ipiv, info = la.dgetrf(A)
    perm = np.arange(m)
    for i, p in enumerate(ipiv):
        perm[[i, p]] = perm[[p, i]]
    assert np.array_equal(pos, perm)
    np.testing.assert_allclose(F, Fl, rtol=0, atol=64 * max(m, n) * EPS * np.abs(A).max() * 10)
    k = min(m, n)
    L = np.tril(F[:, :k], -1) + np.eye(m, k)
    U = np.triu(F[:k])
    assert np.linalg.norm(L @ U - A[pos]) < 16 * max(m, n) * EPS * np.linalg.norm(A)


def test_lu_without_pivoting_and_zero_pivot(orc):
    rng = np.random.default_rng(3)
    A = rng.standard_normal((50, 50)) + 60 * np.eye(50)   # diagonally dominant
    F, pos, done = orc.lu(A, pivot=False)
    assert done == 50 and np.array_equal(pos, np.arange(50))
    L = np.tril(F, -1) + np.eye(50)
    assert np.linalg.norm(L @ np.triu(F) - A) < 1e-13 * np.linalg.norm(A)
    Z = A.copy()
    Z[7, 7] = 0.0
    Z[7, :7] = 0.0      # a zero pivot at column 7 without pivoting
    _, _, done = orc.lu(Z, pivot=False)
    assert done == 7


@pytest.mark.parametrize("n,nrhs", [(1, 1), (7, 3), (100, 1), (250, 6)])
def test_solve_direct_matches_scipy(orc, n, nrhs):
    rng = np.random.default_rng(n + nrhs)
    A = rng.standard_normal((n, n))
    B = rng.standard_normal((n, nrhs))
    X = orc.solve_direct(A, B)
    ref = sla.solve(A, B)
    np.testing.assert_allclose(X, ref, rtol=0, atol=1e-10 * np.abs(ref).max() * n)
    # a planted mistake (transposed system) is caught
    assert not np.allclose(orc.solve_direct(A.T, B), ref, atol=1e-6) or n == 1
