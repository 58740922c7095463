"""GPU parity of the QR adjoint (mqr_qr_backward, FP64 and FP32, strided
batches) against the analytic formula evaluated in numpy FP64
(oracle.qr_adjoint), central finite differences through the
torch.autograd.Function harness (torch.autograd.gradcheck, FP64), and
torch.linalg.qr's own backward (same dgeqrf sign convention for m > n)."""
import numpy as np
import pytest

import mat_amd as M
from oracle.oracle import qr_adjoint

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


def problem(batch, m, n, seed):
    rng = np.random.default_rng(seed)
    Qs, Rs, dQs, dRs = [], [], [], []
    for _ in range(batch):
        A = rng.standard_normal((m, n))
        Q, R = np.linalg.qr(A)
        Qs.append(Q)
        Rs.append(R)
        dQs.append(rng.standard_normal((m, n)))
        dRs.append(np.triu(rng.standard_normal((n, n))))
    return [np.array(x) for x in (Qs, Rs, dQs, dRs)]


def to_dev(torch, x, dt):
    return M.colmajor_batched(torch.from_numpy(x).to(device="cuda", dtype=dt))


@pytest.mark.parametrize("dtype,batch,m,n", [("f64", 3, 60, 20), ("f64", 2, 300, 130),
                                             ("f64", 1, 64, 64), ("f32", 4, 200, 70),
                                             ("f32", 2, 1024, 256), ("f64", 2, 1024, 256)])
def test_adjoint_matches_analytic(h, torch_mod, dtype, batch, m, n):
    torch = torch_mod
    dt = torch.float64 if dtype == "f64" else torch.float32
    note(f"adjoint {dtype} batch={batch} {m}x{n}")
    Q, R, dQ, dR = problem(batch, m, n, seed=m + n)
    dA = M.qr_backward(h, to_dev(torch, Q, dt), to_dev(torch, R, dt), to_dev(torch, dQ, dt),
                       to_dev(torch, dR, dt))
    torch.cuda.synchronize()
    got = dA.cpu().double().numpy()
    u = np.finfo(np.float64 if dtype == "f64" else np.float32).eps
    for i in range(batch):
        Qi, Ri = Q[i], R[i]
        if dtype == "f32":  # the kernel sees the rounded inputs
            Qi, Ri = Qi.astype(np.float32).astype(float), Ri.astype(np.float32).astype(float)
            dQi = dQ[i].astype(np.float32).astype(float)
            dRi = dR[i].astype(np.float32).astype(float)
        else:
            dQi, dRi = dQ[i], dR[i]
        ref = qr_adjoint(Qi, Ri, dQi, dRi)
        kappa = np.linalg.cond(Ri)
        err = np.linalg.norm(got[i] - ref) / np.linalg.norm(ref)
        note(f"  matrix {i}: rel err {err:.2e} (kappa(R) {kappa:.1f})")
        assert err <= 50 * n * kappa * u


def test_adjoint_null_gradients(h, torch_mod):
    """dQ or dR omitted (NULL) counts as zero."""
    torch = torch_mod
    Q, R, dQ, dR = problem(2, 80, 30, seed=5)
    dev = {k: to_dev(torch, v, torch.float64) for k, v in dict(Q=Q, R=R, dQ=dQ, dR=dR).items()}
    a = M.qr_backward(h, dev["Q"], dev["R"], None, dev["dR"]).cpu().numpy()
    b = M.qr_backward(h, dev["Q"], dev["R"], dev["dQ"], None).cpu().numpy()
    for i in range(2):
        assert np.allclose(a[i], qr_adjoint(Q[i], R[i], np.zeros_like(dQ[i]), dR[i]),
                           rtol=1e-10, atol=1e-12)
        assert np.allclose(b[i], qr_adjoint(Q[i], R[i], dQ[i], np.zeros_like(dR[i])),
                           rtol=1e-10, atol=1e-12)


def test_adjoint_gradcheck_finite_differences(torch_mod):
    """torch.autograd.gradcheck: central finite differences of the library's
    own forward QR against the adjoint kernel, FP64."""
    torch = torch_mod
    g = torch.Generator(device="cuda").manual_seed(3)
    A = torch.randn(2, 14, 6, device="cuda", dtype=torch.float64, generator=g)
    A = M.colmajor_batched(A).requires_grad_(True)
    assert torch.autograd.gradcheck(lambda X: M.qr_autograd(X), (A,), eps=1e-6, atol=1e-6,
                                    rtol=1e-5)


@pytest.mark.parametrize("dtype", ["f64", "f32"])
def test_adjoint_full_c5_batch(h, torch_mod, dtype):
    """The whole C5 workload, 256 strided matrices of 1024 x 256, in one
    call; sampled matrices (first, last and two inside) against the analytic
    formula, so that a wrong batch stride or a matrix left out shows."""
    torch = torch_mod
    dt = torch.float64 if dtype == "f64" else torch.float32
    batch, m, n = 256, 1024, 256
    note(f"adjoint full C5 batch {dtype}")
    g = torch.Generator(device="cuda").manual_seed(77)
    A = torch.randn(batch, m, n, generator=g, device="cuda", dtype=torch.float64)
    Q, R = torch.linalg.qr(A.cpu(), mode="reduced")
    dQ = torch.randn(batch, m, n, generator=g, device="cuda", dtype=torch.float64).cpu()
    dR = torch.triu(torch.randn(batch, n, n, generator=g, device="cuda", dtype=torch.float64)).cpu()
    dev = [M.colmajor_batched(x.to(device="cuda", dtype=dt)) for x in (Q, R, dQ, dR)]
    dA = M.qr_backward(h, *dev)
    torch.cuda.synchronize()
    u = np.finfo(np.float64 if dtype == "f64" else np.float32).eps
    for i in (0, 97, 200, batch - 1):
        got = dA[i].cpu().double().numpy()
        Qi, Ri, dQi, dRi = (x[i].to(dt).double().numpy() for x in (Q, R, dQ, dR))
        ref = qr_adjoint(Qi, Ri, dQi, dRi)
        err = np.linalg.norm(got - ref) / np.linalg.norm(ref)
        kappa = np.linalg.cond(Ri)
        note(f"  matrix {i}: rel err {err:.2e} (kappa(R) {kappa:.1f})")
        assert err <= 50 * n * kappa * u


def test_adjoint_gradcheck_larger_fast_mode(torch_mod):
    """gradcheck's fast mode (random projections of the Jacobian) at
    2 x (256 x 64), FP64: the blocked solve's 64-column diagonal blocks and
    the off-diagonal GEMMs all take part."""
    torch = torch_mod
    g = torch.Generator(device="cuda").manual_seed(4)
    A = torch.randn(2, 256, 64, device="cuda", dtype=torch.float64, generator=g)
    A = M.colmajor_batched(A).requires_grad_(True)
    assert torch.autograd.gradcheck(lambda X: M.qr_autograd(X), (A,), eps=1e-6, atol=1e-6,
                                    rtol=1e-5, fast_mode=True)


def test_adjoint_matches_torch_linalg_qr(torch_mod):
    """Loss sum(Wq o Q) + sum(Wr o R) (the C5 workload's form): the gradient
    through the library equals torch.linalg.qr's backward in FP64."""
    torch = torch_mod
    g = torch.Generator(device="cuda").manual_seed(9)
    A = torch.randn(3, 90, 40, device="cuda", dtype=torch.float64, generator=g)
    Wq = torch.randn(3, 90, 40, device="cuda", dtype=torch.float64, generator=g)
    Wr = torch.randn(3, 40, 40, device="cuda", dtype=torch.float64, generator=g)
    A1 = M.colmajor_batched(A.clone()).requires_grad_(True)
    Q1, R1 = M.qr_autograd(A1)
    ((Wq * Q1).sum() + (Wr * R1).sum()).backward()
    A2 = A.detach().cpu().clone().requires_grad_(True)
    Q2, R2 = torch.linalg.qr(A2, mode="reduced")
    assert torch.allclose(Q1.detach().cpu(), Q2.detach(), atol=1e-12)
    ((Wq.cpu() * Q2).sum() + (Wr.cpu() * R2).sum()).backward()
    assert torch.allclose(A1.grad.cpu(), A2.grad, rtol=1e-9, atol=1e-10)
