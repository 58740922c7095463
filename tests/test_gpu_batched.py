"""GPU parity of the batched uneven QR (one persistent launch, FP64 and
FP32) through the C ABI.  Each matrix is checked against the oracle's blocked
Householder factorisation (a restatement of QR/__Householder_Blocked.h:62-245):
R, the reflectors and tau agree to c * m * u * ||A||, with the oracle in FP64
for both precisions (u the precision of the data)."""
import numpy as np
import pytest

import mat_amd as M
from oracle.oracle import Oracle

pytestmark = pytest.mark.gpu


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


def uneven_batch(count, seed, mmax=512, mmin=16):
    rng = np.random.default_rng(seed)
    shapes = []
    for _ in range(count):
        m = int(rng.integers(mmin, mmax + 1))
        n = int(rng.integers(min(8, m), m + 1))
        shapes.append((m, n))
    return shapes, [np.asfortranarray(rng.standard_normal(s)) for s in shapes]


def check_against_oracle(orc, A, VR, tau, eps, c=64.0):
    ref, _, tref = orc.geqrf_blocked(A)
    m, n = A.shape
    k = min(m, n)
    scale = c * max(m, n) * eps * max(np.linalg.norm(A), 1e-300)
    err_vr = np.max(np.abs(VR - ref)) if VR.size else 0.0
    err_tau = np.max(np.abs(tau[:k] - tref)) if k else 0.0
    return err_vr <= scale and err_tau <= c * max(m, n) * eps * 2, (err_vr, scale, err_tau)


@pytest.mark.parametrize("dtype,tile", [("f64", -1), ("f64", 0), ("f32", -1), ("f32", 1)])
def test_batched_matches_oracle(h, orc, torch_mod, dtype, tile):
    torch = torch_mod
    td = torch.float64 if dtype == "f64" else torch.float32
    eps = np.finfo(np.float64 if dtype == "f64" else np.float32).eps
    shapes, mats = uneven_batch(160, seed=11 + tile)
    # edge shapes: single element, one column, wide, square maximum, n = 0
    extra = [(1, 1), (16, 1), (37, 60), (512, 512), (20, 0), (2, 2)]
    rng = np.random.default_rng(5)
    for s in extra:
        shapes.append(s)
        mats.append(np.asfortranarray(rng.standard_normal(s)))
    note(f"batched {dtype} tile={tile}: {len(mats)} matrices")
    dev = [M.colmajor(A, dtype=td) for A in mats]
    taus = M.geqrf_batched(h, dev, tile=tile)
    torch.cuda.synchronize()
    bad = []
    for i, (A, d, t) in enumerate(zip(mats, dev, taus)):
        A_in = A.astype(np.float32).astype(np.float64) if dtype == "f32" else A
        ok, info = check_against_oracle(orc, A_in, np.asfortranarray(d.cpu().double().numpy()),
                                        t.cpu().double().numpy(), eps)
        if not ok:
            bad.append((i, A.shape, info))
    assert not bad, bad[:5]


@pytest.mark.parametrize("dtype", ["f64", "f32"])
def test_batched_layout_boundaries(h, orc, torch_mod, dtype):
    """Shapes at the kernel's internal boundaries, against the oracle: a
    panel's rows split between two register rows per thread at row 256, the
    trailing chunk split into per-wave slices of up to 128 rows (16-row tiles,
    the last one partial), and last panels narrower than 16 columns."""
    torch = torch_mod
    td = torch.float64 if dtype == "f64" else torch.float32
    eps = np.finfo(np.float64 if dtype == "f64" else np.float32).eps
    shapes = [(256, 256), (257, 100), (255, 255), (511, 511), (512, 17), (300, 16), (33, 33),
              (64, 64), (65, 49), (129, 129), (384, 200), (17, 17), (240, 241), (500, 31)]
    rng = np.random.default_rng(17)
    mats = [np.asfortranarray(rng.standard_normal(s)) for s in shapes]
    note(f"batched boundaries {dtype}: {shapes}")
    dev = [M.colmajor(A, dtype=td) for A in mats]
    taus = M.geqrf_batched(h, dev)
    torch.cuda.synchronize()
    bad = []
    for A, d, t in zip(mats, dev, taus):
        A_in = A.astype(np.float32).astype(np.float64) if dtype == "f32" else A
        ok, info = check_against_oracle(orc, A_in, np.asfortranarray(d.cpu().double().numpy()),
                                        t.cpu().double().numpy(), eps)
        if not ok:
            bad.append((A.shape, info))
    assert not bad, bad


def test_batched_special_columns(h, orc, torch_mod):
    """Zero columns keep tau = 0 and identity gives -I exactly (the
    reference's reflector conventions)."""
    rng = np.random.default_rng(3)
    Z = np.asfortranarray(rng.standard_normal((50, 20)))
    Z[:, 4] = 0.0
    Z[:, 11] = 0.0
    I = np.asfortranarray(np.eye(40))
    dev = [M.colmajor(Z), M.colmajor(I)]
    taus = M.geqrf_batched(h, dev)
    torch_mod.cuda.synchronize()
    ok, info = check_against_oracle(orc, Z, np.asfortranarray(dev[0].cpu().numpy()),
                                    taus[0].cpu().numpy(), np.finfo(float).eps)
    assert ok, info
    R = np.triu(dev[1].cpu().numpy())
    assert np.array_equal(R, -np.eye(40))
    assert np.all(taus[1].cpu().numpy() == 2.0)


def test_batched_plan_replays(h, orc, torch_mod):
    """The hipGraph plan replays: two runs on fresh copies of the inputs give
    the same factors as the one-shot call (the kernel's work counters reset
    themselves between replays)."""
    torch = torch_mod
    shapes, mats = uneven_batch(300, seed=21, mmax=200)
    dev = [M.colmajor(A) for A in mats]
    plan = M.BatchedPlan(h, dev)
    outs = []
    for rep in range(2):
        for d, A in zip(dev, mats):
            d.copy_(M.colmajor(A))
        plan.run()
        h.synchronize()
        outs.append([d.cpu().numpy().copy() for d in dev])
    plan.close()
    ref = [M.colmajor(A) for A in mats]
    M.geqrf_batched(h, ref)
    torch.cuda.synchronize()
    for a, b, r in zip(outs[0], outs[1], ref):
        assert np.array_equal(a, b)
        assert np.array_equal(a, r.cpu().numpy())


@pytest.mark.parametrize("per_cu,order", [(1, 0), (1, 1), (0, 2), (2, 1)])
def test_batched_schedules_give_the_same_factors(h, torch_mod, per_cu, order):
    """Every schedule the autotuner can pick (residency, queue order, panel
    width 16) factors each matrix with the same arithmetic: bit-identical to
    the default plan, in FP64 and FP32, over more matrices than the grid has
    workgroups."""
    torch = torch_mod
    note(f"batched schedule per_cu={per_cu} order={order}")
    for dt in (torch.float64, torch.float32):
        shapes, mats = uneven_batch(700, seed=31, mmax=300)
        ref = [M.colmajor(A).to(dt) for A in mats]
        plan0 = M.BatchedPlan(h, ref, tile=1)
        plan0.run()
        h.synchronize()
        got = [M.colmajor(A).to(dt) for A in mats]
        plan = M.BatchedPlan(h, got, tile=1, per_cu=per_cu, order=order)
        plan.run()
        h.synchronize()
        for a, b, ta, tb in zip(got, ref, plan.taus, plan0.taus):
            assert torch.equal(a, b) and torch.equal(ta, tb)
        plan.close()
        plan0.close()


def test_batched_argument_errors(h, torch_mod):
    import ctypes as C
    lib = M.lib()
    m = (C.c_int32 * 1)(600)
    n = (C.c_int32 * 1)(10)
    lda = (C.c_int64 * 1)(600)
    p = (C.c_void_p * 1)(1)
    assert lib.mqr_geqrf_batched_uneven(h.h, 0, 1, m, n, lda, p, p, -1) == -4
    m[0] = 20
    lda[0] = 10
    assert lib.mqr_geqrf_batched_uneven(h.h, 0, 1, m, n, lda, p, p, -1) == -6
    lda[0] = 20
    assert lib.mqr_geqrf_batched_uneven(h.h, 0, 1, m, n, lda, p, p, 7) == -9
    assert lib.mqr_geqrf_batched_uneven(h.h, 5, 1, m, n, lda, p, p, -1) == -2
    assert lib.mqr_geqrf_batched_uneven(h.h, 0, 0, None, None, None, None, None, -1) == 0
