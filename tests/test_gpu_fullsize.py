"""GPU parity at the full workload sizes, through the C ABI.

- Batched uneven QR of the 10000-matrix workload (m ~ U{16..512},
  n ~ U{8..m}, seed 4321): far more matrices than the persistent launch has
  workgroups, so every workgroup dequeues many matrices and reuses its LDS;
  every matrix is checked against the oracle's blocked factorisation
  (QR/__Householder_Blocked.h:62-245), and hipGraph replays of the same work
  must reproduce the one-shot factors bit for bit (the queue counters reset
  themselves at the end of each launch).
- 32768 x 1024 iid and graded: panels taller than 16384 rows, where each
  workgroup of the row-chunk kernels loops over several 64-row chunks; R, the
  leading reflectors and tau against the oracle, ||QR - A|| and ||Q^T Q - I||
  on the GPU.
- 8192 x 8192 iid: the same invariants, and R against LAPACK dgeqrf (OpenBLAS
  through SciPy) column by column.
- 16384 x 4096 least squares with A = U diag(logspace(0, -6)) V^T (U, V from
  torch's QR, independent of this library): plain FP64 and FP32+refinement
  reach a Karlson-Walden backward error <= 1e-14, computed exactly from the
  known SVD.

Tolerances: Householder QR is columnwise backward stable, so two FP64
factorisations of the same matrix differ in R(:, j) by O(u ||a_j||) (LAPACK
against the oracle measures 7 u ||a_j|| iid and 22 u ||a_j|| graded at
32768 x 1024); the tests ask for 512 u ||a_j||.
"""
import os
import time
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

import mat_amd as M
from oracle.oracle import Oracle, ls_backward_error  # noqa: F401  (oracle: checker only)

pytestmark = pytest.mark.gpu
EPS = np.finfo(np.float64).eps
WORKERS = max(1, min(16, os.cpu_count() or 1))


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
    return Oracle("auto")


def host(t):
    return np.asfortranarray(t.detach().cpu().numpy())


# ---------------------------------------------------------------------------
# batched: the persistent queue loop
# ---------------------------------------------------------------------------

def c4_shapes(count=10000, seed=4321):
    """m ~ U{16..512}, n ~ U{8..m}: the batched workload's shapes."""
    rng = np.random.default_rng(seed)
    ms = rng.integers(16, 513, count)
    ns = np.array([rng.integers(8, m + 1) for m in ms])
    return [(int(m), int(n)) for m, n in zip(ms, ns)]


def _check_one(args):
    """Worst error of one GPU factorisation against the oracle, relative to
    its tolerance (<= 1 passes): R and V to 64 max(m, n) u ||A||_F, tau to
    128 max(m, n) u."""
    orc, A, VR, tau, eps = args
    ref, _, tref = orc.geqrf_blocked(A)
    m, n = A.shape
    k = min(m, n)
    scale = 64.0 * max(m, n) * eps * max(np.linalg.norm(A), 1e-300)
    e_vr = np.max(np.abs(VR - ref)) / scale if VR.size else 0.0
    e_tau = np.max(np.abs(tau[:k] - tref)) / (128.0 * max(m, n) * eps) if k else 0.0
    return max(e_vr, e_tau)


@pytest.mark.parametrize("dtype", ["f64", "f32"])
def test_batched_persistent_loop_matches_oracle(h, orc, torch_mod, dtype):
    torch = torch_mod
    td = torch.float64 if dtype == "f64" else torch.float32
    eps = np.finfo(np.float64 if dtype == "f64" else np.float32).eps
    shapes = c4_shapes()
    # the launch has at most 3 workgroups per CU (FP32; 2 in FP64): with
    # 10000 matrices every workgroup factors a dozen or more of them
    assert len(shapes) >= 3 * 3 * 256
    rng = np.random.default_rng(99)
    mats = [np.asfortranarray(rng.standard_normal(s)) for s in shapes]
    if dtype == "f32":
        mats = [np.asfortranarray(A.astype(np.float32).astype(np.float64)) for A in mats]
    note(f"batched persistent loop {dtype}: {len(mats)} matrices, one-shot")
    dev = [M.colmajor(A, dtype=td) for A in mats]
    taus = M.geqrf_batched(h, dev)
    torch.cuda.synchronize()
    got = [(np.asfortranarray(d.cpu().double().numpy()), t.cpu().double().numpy())
           for d, t in zip(dev, taus)]
    note(f"batched persistent loop {dtype}: oracle check on {WORKERS} threads")
    with ThreadPoolExecutor(WORKERS) as ex:
        worst = list(ex.map(_check_one, [(orc, A, g[0], g[1], eps) for A, g in zip(mats, got)]))
    bad = [(i, shapes[i], w) for i, w in enumerate(worst) if not w <= 1.0]
    note(f"batched persistent loop {dtype}: worst error / tolerance {max(worst):.3g}")
    assert not bad, bad[:5]

    # hipGraph plan: two replays over restored inputs, bit-identical to the
    # one-shot factors
    note(f"batched persistent loop {dtype}: plan replays")
    work = [M.colmajor(A, dtype=td) for A in mats]
    plan = M.BatchedPlan(h, work)
    for rep in range(2):
        for w, A in zip(work, mats):
            w.copy_(torch.from_numpy(np.ascontiguousarray(A.T)).to(td).T)
        plan.run()
        h.synchronize()
        for i, (w, t, g) in enumerate(zip(work, plan.taus, got)):
            assert np.array_equal(np.asfortranarray(w.cpu().double().numpy()), g[0]), (rep, i)
            assert np.array_equal(t.cpu().double().numpy(), g[1]), (rep, i)
    plan.close()


# ---------------------------------------------------------------------------
# tall panels past 16384 rows, and 8192^2
# ---------------------------------------------------------------------------

def haar_torch(torch, rows, cols, seed):
    """Orthonormal columns from torch's own QR (independent of this
    library), signs fixed by diag(R)."""
    g = torch.Generator(device="cuda").manual_seed(seed)
    G = torch.randn(rows, cols, generator=g, device="cuda", dtype=torch.float64)
    Q, R = torch.linalg.qr(G)
    d = torch.sign(torch.diagonal(R))
    d[d == 0] = 1.0
    return Q * d[None, :]


def graded(torch, m, n, smin, seed):
    U = haar_torch(torch, m, n, seed)
    V = haar_torch(torch, n, n, seed + 1)
    s = torch.logspace(0.0, float(np.log10(smin)), n, device="cuda", dtype=torch.float64)
    return (U * s[None, :]) @ V.T, U, s


def factor_and_check(torch, h, A0):
    """Factor a copy of the column-major device matrix A0 on the GPU, form
    the thin Q, and return (VR, tau, ||QR - A||/||A||, ||Q^T Q - I||)."""
    m, n = A0.shape
    dA = M.empty_colmajor(m, n)
    dA.copy_(A0)
    tau, T = M.geqrf(h, dA)
    Q = M.orgqr(h, dA, T, n, nq=n)
    h.synchronize()
    R = torch.triu(dA[:n, :])
    res = float(torch.linalg.norm(Q @ R - A0) / torch.linalg.norm(A0))
    orth = float(torch.linalg.norm(Q.T @ Q - torch.eye(n, device="cuda", dtype=torch.float64)))
    VR = host(dA)
    tau = tau.cpu().numpy()
    del Q, R, dA, T
    torch.cuda.empty_cache()
    return VR, tau, res, orth


@pytest.mark.parametrize("kind", ["iid", "graded"])
def test_tall_32768x1024_matches_oracle(h, orc, torch_mod, kind):
    torch = torch_mod
    m, n = 32768, 1024
    note(f"tall {m}x{n} {kind}")
    if kind == "iid":
        g = torch.Generator(device="cuda").manual_seed(31)
        A0 = M.colmajor(torch.randn(m, n, generator=g, device="cuda", dtype=torch.float64))
    else:
        A, _, _ = graded(torch, m, n, 1e-8, 32)
        A0 = M.colmajor(A)
        del A
    VR, tau, res, orth = factor_and_check(torch, h, A0)
    note(f"tall {kind}: |QR-A|/|A| {res:.2e} |QtQ-I| {orth:.2e}")
    assert res <= 8 * m * EPS and orth <= 8 * m * EPS
    A = host(A0)
    note(f"tall {kind}: oracle factorisation")
    VRo, _, tau_o = orc.geqrf_blocked(A)
    colnorm = np.linalg.norm(A, axis=0)
    dR = np.abs(np.triu(VR)[:n] - np.triu(VRo)[:n]) / colnorm[None, :]
    note(f"tall {kind}: max |dR_ij| / (u |a_j|) = {dR.max() / EPS:.1f}")
    assert dR.max() <= 512 * EPS
    # the leading quarter of the reflectors is well determined in both cases
    lead = n // 4
    dV = np.abs(np.tril(VR, -1)[:, :lead] - np.tril(VRo, -1)[:, :lead]).max()
    dT = np.abs(tau[:lead] - tau_o[:lead]).max()
    note(f"tall {kind}: leading V {dV:.2e} tau {dT:.2e}")
    assert dV <= 512 * EPS and dT <= 512 * EPS


@pytest.mark.parametrize("m", [16384, 20000])
def test_breakdown_panels_at_full_height(h, torch_mod, m):
    """Every 64-column panel holds duplicated and exactly zero columns, so
    the CholeskyQR panel breaks down each time and the panel is factored
    column by column (QR/__Householder_Blocked.h:62-84) by all workgroups of
    the row-chunk kernel; at 20000 rows each of them holds two 64-row chunks.
    On the GPU: ||QR - A||, ||Q^T Q - I||; the reference's convention for a
    zero column (tau = 0, the column left as it is); and a time that only
    the many-workgroup factorisation reaches (one workgroup per panel took
    about 46 ms per 16384-row panel, 0.74 s for these 16 panels)."""
    torch = torch_mod
    n = 1024
    note(f"breakdown panels {m}x{n}")
    g = torch.Generator(device="cuda").manual_seed(m)
    A = torch.randn(m, n, generator=g, device="cuda", dtype=torch.float64)
    A[:, 1::8] = A[:, 0::8]   # duplicated columns in every panel
    A[:, 5::64] = 0.0         # one zero column per panel
    A0 = M.colmajor(A)
    del A
    dA = M.empty_colmajor(m, n)
    best = float("inf")
    for _ in range(2):
        dA.copy_(A0)
        h.synchronize()
        t0 = time.perf_counter()
        M.geqrf(h, dA)
        h.synchronize()
        best = min(best, time.perf_counter() - t0)
    del dA
    note(f"breakdown panels {m}x{n}: factorisation {best * 1e3:.1f} ms")
    assert best < 0.3
    VR, tau, res, orth = factor_and_check(torch, h, A0)
    note(f"breakdown panels: |QR-A|/|A| {res:.2e} |QtQ-I| {orth:.2e}")
    assert res <= 8 * m * EPS and orth <= 8 * m * EPS
    zc = np.arange(5, n, 64)
    assert np.all(tau[zc] == 0.0)
    for j in zc:
        assert np.all(VR[j:, j] == 0.0)
    # a duplicated column leaves a pivot at rounding level
    d = np.abs(np.diag(VR[:n]))
    colnorm = np.linalg.norm(host(A0), axis=0)
    assert np.all(d[1::8] <= 1e-12 * colnorm[1::8])


def test_square_8192_invariants_and_lapack_r(h, torch_mod):
    torch = torch_mod
    m = n = 8192
    note(f"square {m}x{n} iid")
    g = torch.Generator(device="cuda").manual_seed(41)
    A0 = M.colmajor(torch.randn(m, n, generator=g, device="cuda", dtype=torch.float64))
    VR, tau, res, orth = factor_and_check(torch, h, A0)
    note(f"square: |QR-A|/|A| {res:.2e} |QtQ-I| {orth:.2e}")
    assert res <= 8 * m * EPS and orth <= 8 * m * EPS
    import scipy.linalg.lapack as la
    A = host(A0)
    note("square: LAPACK dgeqrf")
    qr, tau_l, _, info = la.dgeqrf(A.copy(order="F"))
    assert info == 0
    R, Rl = np.triu(VR), np.triu(qr)
    # the last column has nothing below the diagonal: the reference reflects
    # it (tau = 2, R_nn = -a_nn) where LAPACK leaves it (tau = 0); the GPU's
    # last 64-row panel forms tau from its reconstruction, so 2 up to rounding
    assert abs(tau[-1] - 2.0) <= 512 * EPS and tau_l[-1] == 0.0
    Rl[-1, -1] = -Rl[-1, -1]
    colnorm = np.linalg.norm(A, axis=0)
    dR = np.abs(R - Rl) / colnorm[None, :]
    note(f"square: max |dR_ij| / (u |a_j|) = {dR.max() / EPS:.1f}")
    assert dR.max() <= 512 * EPS
    # tau of the short trailing columns follows the rounding of the trailing
    # submatrix more closely than R does: the leading half to 512 u, the rest
    # to 1e-11
    np.testing.assert_allclose(tau[:n // 2], tau_l[:n // 2], rtol=0, atol=512 * EPS)
    np.testing.assert_allclose(tau[:-1], tau_l[:-1], rtol=0, atol=1e-11)


# ---------------------------------------------------------------------------
# the 16384 x 4096 least-squares problem
# ---------------------------------------------------------------------------

def kw_backward_error(torch, A, b, x, U, s):
    """Karlson-Walden ||(A^T A + phi^2 I)^{-1/2} A^T r|| / (||x|| ||A||_F),
    phi = ||r|| / ||x||, from the known SVD A = U diag(s) V^T (V drops out:
    A^T r = V diag(s) U^T r and V is orthogonal)."""
    r = b - A @ x
    nx, nr, nA = torch.linalg.norm(x), torch.linalg.norm(r), torch.linalg.norm(A)
    phi = nr / nx
    y = s * (U.T @ r)
    return float(torch.linalg.norm(y / torch.sqrt(s * s + phi * phi)) / nx / nA)


@pytest.fixture(scope="module")
def ls_problem(torch_mod):
    torch = torch_mod
    m, n = 16384, 4096
    A, U, s = graded(torch, m, n, 1e-6, 1234)
    A = M.colmajor(A)
    g = torch.Generator(device="cuda").manual_seed(1236)
    x0 = torch.randn(n, generator=g, device="cuda", dtype=torch.float64)
    b = A @ x0 + 1e-10 * torch.randn(m, generator=g, device="cuda", dtype=torch.float64)
    yield A, b.contiguous(), U, s
    del A, U, s
    torch.cuda.empty_cache()


def test_ls_16384x4096_plain_backward_error(h, torch_mod, ls_problem):
    torch = torch_mod
    A, b, U, s = ls_problem
    m, n = A.shape
    note(f"least squares {m}x{n} plain FP64")
    Aw = M.empty_colmajor(m, n)
    Aw.copy_(A)
    bw = M.empty_colmajor(m, 1)
    bw[:, 0].copy_(b)
    x = M.gels(h, Aw, bw)[:, 0].clone()
    h.synchronize()
    del Aw
    kw = kw_backward_error(torch, A, b, x, U, s)
    note(f"least squares plain: KW backward error {kw:.2e}")
    assert kw <= 1e-14


@pytest.mark.parametrize("prec", ["fp32"])
def test_ls_16384x4096_mixed_backward_error(h, torch_mod, ls_problem, prec):
    torch = torch_mod
    A, b, U, s = ls_problem
    note(f"least squares {A.shape[0]}x{A.shape[1]} mixed {prec}")
    x, info = M.gels_mixed(h, prec, A, b)
    h.synchronize()
    kw = kw_backward_error(torch, A, b, x, U, s)
    note(f"least squares mixed {prec}: KW {kw:.2e}, {info.iterations} steps, "
         f"converged {info.converged}, fell back {info.fell_back}")
    assert info.converged == 1 and info.fell_back == 0
    assert kw <= 1e-14


def test_dgels_many_right_hand_sides(h, torch_mod):
    """More right-hand sides than the factorisation's partial-product buffer
    holds at once (100 x 64 with 20000 columns): B is applied in column
    chunks; X against numpy's lstsq."""
    rng = np.random.default_rng(77)
    m, n, nrhs = 100, 64, 20000
    A = rng.standard_normal((m, n))
    B = rng.standard_normal((m, nrhs))
    note(f"dgels {m}x{n} nrhs={nrhs}")
    dA, dB = M.colmajor(A), M.colmajor(B)
    X = host(M.gels(h, dA, dB))
    h.synchronize()
    Xref = np.linalg.lstsq(A, B, rcond=None)[0]
    np.testing.assert_allclose(X, Xref, rtol=0, atol=1e-11 * max(1.0, np.abs(Xref).max()))
