"""GPU parity of the FP32 GEMM tile and of the mixed-precision least-squares
solve (FP32 trailing updates, FP64 panels, FP64 refinement), through the C ABI.

The solve has no counterpart in the reference; it is checked against the
oracle's FP64 least-squares solution (a restatement of the reference's blocked
Householder QR followed by Q^T b and back substitution) and against the exact
normwise backward error computed from the SVD (Karlson-Walden formula,
oracle.ls_backward_error): the target is the FP64 level, backward error
<= 1e-14 after refinement."""
import numpy as np
import pytest

import mat_amd as M
from oracle.oracle import Oracle, ls_backward_error

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


def ls_problem(m, n, cond, seed, resid=1e-10):
    """A = U diag(logspace(0, -log10 cond)) V^T with random orthogonal U, V;
    b = A x0 + resid * (unit-norm noise) * ||A x0||."""
    rng = np.random.default_rng(seed)
    U, _ = np.linalg.qr(rng.standard_normal((m, n)))
    V, _ = np.linalg.qr(rng.standard_normal((n, n)))
    s = np.logspace(0, -np.log10(cond), n)
    A = np.asfortranarray((U * s) @ V.T)
    x0 = rng.standard_normal(n)
    b0 = A @ x0
    e = rng.standard_normal(m)
    b = b0 + resid * np.linalg.norm(b0) * e / np.linalg.norm(e)
    return A, b


# ---------------------------------------------------------------------------
# FP32 MFMA GEMM tile
# ---------------------------------------------------------------------------

@pytest.mark.parametrize("ta,tb", [("N", "N"), ("T", "N"), ("N", "T"), ("T", "T")])
def test_sgemm_exact_on_integers(h, torch_mod, ta, tb):
    """Integer inputs with |sums| < 2^24: every FP32 product and sum is exact,
    so the result must equal the integer product exactly (this pins the MFMA
    accumulator layout and every edge mask)."""
    rng = np.random.default_rng(7)
    for (m, n, k) in [(1, 1, 1), (16, 16, 4), (64, 64, 16), (37, 29, 41), (200, 150, 90),
                      (1000, 70, 300), (513, 384, 200), (128, 128, 16), (260, 1030, 33)]:
        note(f"sgemm {ta}{tb} m={m} n={n} k={k}")
        A = rng.integers(-8, 9, (k, m) if ta == "T" else (m, k)).astype(np.float32)
        B = rng.integers(-8, 9, (n, k) if tb == "T" else (k, n)).astype(np.float32)
        Cm = rng.integers(-8, 9, (m, n)).astype(np.float32)
        opA = A.T if ta == "T" else A
        opB = B.T if tb == "T" else B
        exact = 2 * Cm.astype(np.int64) - opA.astype(np.int64) @ opB.astype(np.int64)
        f32 = torch_mod.float32
        dC = M.colmajor(Cm, dtype=f32)
        M.gemm(h, ta, tb, -1.0, M.colmajor(A, dtype=f32), M.colmajor(B, dtype=f32), 2.0, dC)
        got = host(dC)
        assert np.array_equal(got.astype(np.int64), exact), (ta, tb, m, n, k)


def test_sgemm_random_against_fp64(h, torch_mod):
    rng = np.random.default_rng(8)
    m, n, k = 700, 900, 1500
    A = rng.standard_normal((m, k)).astype(np.float32)
    B = rng.standard_normal((k, n)).astype(np.float32)
    ref = A.astype(np.float64) @ B.astype(np.float64)
    f32 = torch_mod.float32
    dC = M.empty_colmajor(m, n, dtype=f32)
    M.gemm(h, "N", "N", 1.0, M.colmajor(A, dtype=f32), M.colmajor(B, dtype=f32), 0.0, dC)
    got = host(dC).astype(np.float64)
    # FP32 accumulation over k terms: |err| <= k u32 sum |a||b|
    bound = k * np.finfo(np.float32).eps * (np.abs(A).astype(np.float64) @ np.abs(B))
    assert np.all(np.abs(got - ref) <= bound)


# ---------------------------------------------------------------------------
# mixed-precision least squares
# ---------------------------------------------------------------------------

def run_mixed(h, A, b, prec="fp32", **kw):
    dA = M.colmajor(A)
    import torch
    db = torch.from_numpy(b).cuda()
    x, info = M.gels_mixed(h, prec, dA, db, **kw)
    h.synchronize()
    # A and b are inputs only
    assert np.array_equal(host(dA), A)
    assert np.array_equal(db.cpu().numpy(), b)
    return x.cpu().numpy(), info


@pytest.mark.parametrize("m,n,cond,resid", [(2000, 500, 1e6, 1e-10), (1000, 333, 1e2, 1e-10),
                                            (1500, 300, 1e4, 1.0), (700, 700, 1e3, 0.0),
                                            (4100, 130, 1e5, 1e-6), (1003, 257, 1e4, 1e-8)])
def test_mixed_fp32_reaches_fp64_backward_error(h, orc, m, n, cond, resid):
    note(f"gels_mixed fp32 m={m} n={n} cond={cond:g} resid={resid:g}")
    A, b = ls_problem(m, n, cond, seed=m + n, resid=resid)
    x, info = run_mixed(h, A, b)
    note(f"  iterations={info.iterations} converged={info.converged} "
         f"fell_back={info.fell_back} omega={info.backward_error:.3e}")
    assert info.converged == 1 and info.fell_back == 0
    be = ls_backward_error(A, b, x)
    assert be["kw"] <= 1e-14, be
    # the reported estimate bounds the exact backward error from above
    assert info.backward_error <= 1e-14
    assert be["kw"] <= 1.5 * info.backward_error + 1e-17
    assert abs(info.residual_norm - be["resid"]) <= 1e-6 * max(be["resid"], 1e-300) + 1e-12
    xo = orc.gels(A, b)
    tol = 50 * np.sqrt(n) * cond * EPS * (1.0 + cond * resid)
    assert np.linalg.norm(x - xo) <= tol * np.linalg.norm(xo)


def one_pass_estimates(A, nb=64):
    """The estimate on which the mixed factorisation skips the second
    CholeskyQR pass of a 64-column panel j >= 1 (one pass when <= 1e-9):
    u ||R1^{-1}||_F^2 (||R1||_F^2 + 2 tr Gold), where R1 is the panel's R
    factor after the earlier panels' reflectors (the diagonal block R_jj of
    A's R) and Gold = C^T C of the panel's columns over the previous panel's
    frame (the rows from the previous panel's top: R[j-1:j+1, j]^T R[j-1:j+1, j]
    in 64-blocks).  Computed here from numpy's QR of A, up to rounding."""
    R = np.linalg.qr(A, mode="r")
    n = A.shape[1]
    est = []
    for j in range(nb, n - nb + 1, nb):
        R1 = R[j:j + nb, j:j + nb]
        gold = np.sum(R[j - nb:j + nb, j:j + nb] ** 2)
        ri = np.linalg.inv(R1)
        est.append(2.0 ** -53 * np.sum(ri ** 2) * (np.sum(R1 ** 2) + 2.0 * gold))
    return np.array(est)


@pytest.mark.parametrize("cond,panel_cond", [(1e3, 1.0), (10.0, 1e4)])
def test_mixed_one_and_two_pass_panels(h, orc, cond, panel_cond):
    """Well-conditioned panels take one CholeskyQR pass in the mixed
    factorisation; panels whose columns are scaled over panel_cond = 1e4 take
    the second pass.  The construction lands on the intended side of the
    device's test (checked on the CPU), and both converge to the FP64
    backward-error level without the fallback."""
    m, n = 3000, 640
    note(f"gels_mixed fp32 panel scaling {panel_cond:g}")
    A, _ = ls_problem(m, n, cond, seed=77)
    A = np.asfortranarray(A * np.tile(np.logspace(0, -np.log10(panel_cond), 64), n // 64))
    rng = np.random.default_rng(78)
    b0 = A @ rng.standard_normal(n)
    e = rng.standard_normal(m)
    b = b0 + 1e-10 * np.linalg.norm(b0) * e / np.linalg.norm(e)
    est = one_pass_estimates(A)
    if panel_cond > 1.0:
        assert est.min() > 1e-7, est
    else:
        assert est.max() < 1e-10, est
    x, info = run_mixed(h, A, b)
    note(f"  iterations={info.iterations} converged={info.converged} "
         f"fell_back={info.fell_back} omega={info.backward_error:.3e}")
    assert info.converged == 1 and info.fell_back == 0
    be = ls_backward_error(A, b, x)
    assert be["kw"] <= 1e-14, be
    xo = orc.gels(A, b)
    cond = np.linalg.cond(A)
    tol = 50 * np.sqrt(n) * cond * EPS * (1.0 + cond * 1e-10)
    assert np.linalg.norm(x - xo) <= tol * np.linalg.norm(xo)


def test_mixed_fp64_refinement(h, orc):
    A, b = ls_problem(1200, 400, 1e8, seed=3, resid=1e-8)
    x, info = run_mixed(h, A, b, prec="fp64")
    assert info.converged == 1 and info.iterations <= 3
    be = ls_backward_error(A, b, x)
    assert be["kw"] <= 1e-14, be


def test_mixed_ill_conditioned_falls_back(h, orc):
    """kappa = 1e10 is beyond what an FP32 factorisation can precondition:
    refinement stagnates or diverges and the FP64 solve takes over."""
    A, b = ls_problem(1500, 400, 1e10, seed=5, resid=1e-12)
    x, info = run_mixed(h, A, b)
    assert info.fell_back == 1 and info.converged == 0
    be = ls_backward_error(A, b, x)
    assert be["kw"] <= 1e-14, be
    xo = orc.gels(A, b)
    assert np.linalg.norm(x - xo) <= 1e-3 * np.linalg.norm(xo)


def test_mixed_without_fallback_reports_non_convergence(h):
    A, b = ls_problem(1500, 400, 1e10, seed=5, resid=1e-12)
    with pytest.raises(M.MqrError) as e:
        run_mixed(h, A, b, fallback=False)
    assert e.value.status == M.MQR_ERR_NOT_CONVERGED


def test_mixed_zero_rhs_and_single_column(h):
    A, _ = ls_problem(500, 64, 10.0, seed=9)
    x, info = run_mixed(h, A, np.zeros(500))
    assert info.converged == 1 and np.all(x == 0.0)
    rng = np.random.default_rng(2)
    a = rng.standard_normal((300, 1))
    b = rng.standard_normal(300)
    x, info = run_mixed(h, np.asfortranarray(a), b)
    assert info.converged == 1
    exact = (a[:, 0] @ b) / (a[:, 0] @ a[:, 0])
    assert abs(x[0] - exact) <= 1e-14 * abs(exact)


def test_mixed_repeated_calls_reuse_workspace(h):
    """Two solves in a row, the second smaller: the workspace carved for the
    first must not leak state into the second."""
    A1, b1 = ls_problem(3000, 600, 1e3, seed=11)
    A2, b2 = ls_problem(900, 200, 1e3, seed=12)
    x1, _ = run_mixed(h, A1, b1)
    x2, _ = run_mixed(h, A2, b2)
    assert ls_backward_error(A1, b1, x1)["kw"] <= 1e-14
    assert ls_backward_error(A2, b2, x2)["kw"] <= 1e-14


def test_host_least_squares_api(h):
    A, b = ls_problem(800, 200, 1e4, seed=13)
    x, info = M.LeastSquares(A, b, prec="fp32", handle=h)
    assert info.converged == 1
    assert ls_backward_error(A, b, x)["kw"] <= 1e-14


# ---------------------------------------------------------------------------
# BF16 MFMA products (split-BF16 trailing updates)
# ---------------------------------------------------------------------------

@pytest.mark.parametrize("terms", [1, 3])
@pytest.mark.parametrize("ta,tb", [("N", "N"), ("T", "N"), ("N", "T"), ("T", "T")])
def test_bf16_gemm_exact_on_small_integers(h, torch_mod, terms, ta, tb):
    """|x| <= 8 is exact in BF16 (lo = 0) and every sum stays below 2^24, so
    both modes must return the integer product exactly: this pins the BF16
    MFMA operand and accumulator layouts and the edge masks."""
    rng = np.random.default_rng(17)
    f32 = torch_mod.float32
    for (m, n, k) in [(16, 16, 16), (37, 29, 41), (200, 150, 90), (513, 384, 200), (128, 260, 33)]:
        A = rng.integers(-8, 9, (k, m) if ta == "T" else (m, k)).astype(np.float32)
        B = rng.integers(-8, 9, (n, k) if tb == "T" else (k, n)).astype(np.float32)
        Cm = rng.integers(-8, 9, (m, n)).astype(np.float32)
        opA = A.T if ta == "T" else A
        opB = B.T if tb == "T" else B
        exact = Cm.astype(np.int64) - opA.astype(np.int64) @ opB.astype(np.int64)
        dC = M.colmajor(Cm, dtype=f32)
        M.gemm_bf16(h, terms, ta, tb, -1.0, M.colmajor(A, dtype=f32), M.colmajor(B, dtype=f32),
                    1.0, dC)
        assert np.array_equal(host(dC).astype(np.int64), exact), (terms, ta, tb, m, n, k)


@pytest.mark.parametrize("terms,bits", [(1, 8), (3, 16)])
def test_bf16_gemm_error_bound(h, torch_mod, terms, bits):
    """Random operands: |C - AB| <= k 2^-bits |A||B| (relative error of one
    product 2^-(bits-1), FP32 accumulation on top)."""
    rng = np.random.default_rng(19)
    m, n, k = 300, 280, 700
    A = rng.standard_normal((m, k)).astype(np.float32)
    B = rng.standard_normal((k, n)).astype(np.float32)
    f32 = torch_mod.float32
    dC = M.empty_colmajor(m, n, dtype=f32)
    M.gemm_bf16(h, terms, "N", "N", 1.0, M.colmajor(A, dtype=f32), M.colmajor(B, dtype=f32), 0.0,
                dC)
    ref = A.astype(np.float64) @ B.astype(np.float64)
    bound = (2.0 ** (1 - bits) + k * np.finfo(np.float32).eps) * (np.abs(A).astype(np.float64)
                                                                   @ np.abs(B))
    err = np.abs(host(dC).astype(np.float64) - ref)
    assert np.all(err <= bound)
    # and the split form is markedly better than one BF16 product
    if terms == 3:
        assert np.max(err / np.maximum(bound, 1e-300)) < 0.5


@pytest.mark.parametrize("prec,cond,converges", [("bf16x3", 1e3, True), ("bf16", 10.0, True),
                                                 ("bf16", 1e6, False)])
def test_mixed_bf16_refinement(h, orc, prec, cond, converges):
    """BF16 trailing updates: refinement reaches the FP64 solution while
    u_bf16 * kappa is small; otherwise it stagnates and the FP64 solve runs
    (fell_back), still returning the FP64 least-squares solution."""
    A, b = ls_problem(1500, 400, cond, seed=int(cond) % 97 + 1, resid=1e-8)
    x, info = run_mixed(h, A, b, prec=prec)
    note(f"  {prec} cond={cond:g}: iterations={info.iterations} converged={info.converged} "
         f"fell_back={info.fell_back}")
    assert (info.converged == 1) == converges
    assert info.fell_back == (0 if converges else 1)
    assert ls_backward_error(A, b, x)["kw"] <= 1e-14


def test_mixed_nonfinite_input_falls_back(h):
    """A NaN in A: refinement stops at the first non-finite estimate instead of
    iterating, the FP64 solve takes over, and its result carries the NaN as
    the reference's arithmetic would."""
    import torch
    A, b = ls_problem(600, 150, 1e3, seed=21)
    A[5, 7] = np.nan
    x, info = M.gels_mixed(h, "fp32", M.colmajor(A), torch.from_numpy(b).cuda())
    h.synchronize()
    assert info.converged == 0 and info.fell_back == 1
    assert not np.all(np.isfinite(x.cpu().numpy()))


@pytest.mark.parametrize("nb,prec", [(256, "fp32"), (128, "fp32"), (256, "fp64")])
def test_mixed_outer_blocks(torch_mod, orc, nb, prec):
    """With a block size above 64 the solves use two-level blocking (64-column
    FP64 panels, each nb-column block applied to the far columns as one
    block reflector, in FP32 for the mixed solve): the same FP64-level
    solution, backward error <= 1e-14."""
    m, n = 3000, 900
    A, b = ls_problem(m, n, 1e4, seed=nb + len(prec))
    note(f"outer blocks nb={nb} {prec} {m}x{n}")
    with M.Handle(0, nb=nb) as hn:
        x, info = M.gels_mixed(hn, prec, M.colmajor(A), M.colmajor(b[:, None])[:, 0].contiguous())
        hn.synchronize()
        x = x.cpu().numpy()
    be = ls_backward_error(A, b, x)
    note(f"  iterations={info.iterations} converged={info.converged} kw={be['kw']:.2e}")
    assert info.fell_back == 0 and be["kw"] <= 1e-14
    xo = orc.gels(A, b)
    np.testing.assert_allclose(x, xo, rtol=0, atol=1e-8 * np.abs(xo).max())


@pytest.mark.parametrize("prec,cond", [("bf16x3", 1e6), ("bf16x3", 1e5)])
def test_mixed_krylov_stage_converges(h, orc, prec, cond):
    """Split-BF16 updates at kappa = 1e6: u * kappa > 1, so the stationary
    refinement stalls; conjugate gradients preconditioned by the split-BF16 R
    (condition of A R^{-1} about 1 + u kappa) carry the solution to FP64
    backward error without the FP64 fallback."""
    m, n = 2000, 500
    A, b = ls_problem(m, n, cond, seed=31, resid=1e-10)
    x, info = run_mixed(h, A, b, prec=prec)
    be = ls_backward_error(A, b, x)
    note(f"  {prec} cond={cond:g}: iterations={info.iterations} krylov={info.krylov_iterations} "
         f"converged={info.converged} fell_back={info.fell_back} kw={be['kw']:.2e}")
    assert info.converged == 1 and info.fell_back == 0
    assert be["kw"] <= 1e-14
    xo = orc.gels(A, b)
    np.testing.assert_allclose(x, xo, rtol=0, atol=1e-7 * np.abs(xo).max())


def test_mixed_krylov_stage_can_be_switched_off(h):
    A, b = ls_problem(2000, 500, 1e6, seed=31, resid=1e-10)
    x, info = run_mixed(h, A, b, prec="bf16x3", krylov_max_iter=-1)
    assert info.krylov_iterations == 0 and info.fell_back == 1
    assert ls_backward_error(A, b, x)["kw"] <= 1e-14
