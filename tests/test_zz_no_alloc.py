"""The allocation-free batched entry point (mqr_geqrf_batched_uneven_noalloc)
against the allocating one (mqr_geqrf_batched_uneven), through the C ABI:
two calls in a row on the same handle, the second with a smaller batch, each
bit-identical to the allocating path (the same kernel on the same inputs; the
order in which workgroups take matrices does not change any matrix's
factors).  The batch limit is rejected before any work."""
import ctypes as C

import numpy as np
import pytest

import mat_amd as M

pytestmark = pytest.mark.gpu


def note(msg):
    print(msg, flush=True)


@pytest.fixture(scope="module")
def torch_mod():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a GPU"
    return torch


def uneven(count, seed, mmax):
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(count):
        m = int(rng.integers(1, mmax + 1))
        n = int(rng.integers(0, min(m + 24, 512) + 1))
        out.append(np.asfortranarray(rng.standard_normal((m, n))))
    return out


@pytest.mark.parametrize("dtype", ["f64", "f32"])
def test_noalloc_matches_allocating_path(torch_mod, dtype):
    torch = torch_mod
    td = torch.float64 if dtype == "f64" else torch.float32
    with M.Handle(0) as h:
        for call, (count, seed, mmax) in enumerate([(2500, 1, 512), (700, 2, 300)]):
            mats = uneven(count, seed + (10 if dtype == "f32" else 0), mmax)
            note(f"no-alloc {dtype} call {call}: {count} matrices, m <= {mmax}")
            ref = [M.colmajor(A, dtype=td) for A in mats]
            tau_ref = M.geqrf_batched(h, ref)
            got = [M.colmajor(A, dtype=td) for A in mats]
            tau_got = [torch.full((max(min(t.shape), 1),), float("nan"), device="cuda", dtype=td)
                       for t in got]
            M.geqrf_batched_noalloc(h, got, tau_got)
            h.synchronize()
            for i, (a, b, ta, tb) in enumerate(zip(ref, got, tau_ref, tau_got)):
                k = min(a.shape)
                assert torch.equal(a, b), (call, i, tuple(a.shape))
                assert torch.equal(ta[:k], tb[:k]), (call, i, tuple(a.shape))


def test_noalloc_rejects_batches_beyond_its_scratch(torch_mod):
    with M.Handle(0) as h:
        L = M.lib()
        b = M.NOALLOC_MAX_BATCH + 1
        m = (C.c_int32 * b)(*([4] * b))
        n = (C.c_int32 * b)(*([4] * b))
        lda = (C.c_int64 * b)(*([4] * b))
        p = (C.c_void_p * b)(*([1] * b))
        st = L.mqr_geqrf_batched_uneven_noalloc(h.h, M.MQR_F64, b, m, n, lda, p, p, -1)
        assert st == M.MQR_ERR_TOO_LARGE
        assert "scratch" in M.strerror(st)
        # argument errors keep their LAPACK-style positions
        m[0] = 600
        assert L.mqr_geqrf_batched_uneven_noalloc(h.h, M.MQR_F64, 1, m, n, lda, p, p, -1) == -4
        assert L.mqr_geqrf_batched_uneven_noalloc(h.h, 7, 1, m, n, lda, p, p, -1) == -2
