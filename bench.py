"""Benchmark: QR least squares at 16384 x 4096 on one MI355X.

Workload (the headline configuration): min ||A x - b|| with
A = U diag(s) V^T, s = logspace(0, -6, n), U (m x n) and V (n x n) Haar
orthogonal (Q factor of a seeded Gaussian with diag(R) signs removed),
x0 ~ N(0, 1), b = A x0 + 1e-10 N(0, 1).  Inputs are synthetic, generated on the
device from numpy seeds, and resident in HBM before timing.

A step is one complete solve through the C ABI:
  plain : device copy of A, FP64 blocked Householder QR, Q^T b, R x = c
  mixed : FP32 MFMA trailing updates, FP64 panels / T / residuals, refinement
          on the augmented system to FP64 backward error (A is not modified)
TFLOP/s uses the LAPACK count 2mn^2 - 2n^3/3 + 4mn - n^2 (refinement work is
not counted).  The JSON line reports the mixed path as `value` when it is
enabled (default), with the plain FP64 path beside it at its own error.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "QR TFLOP/s at 16384x4096: mixed+IR vs FP64 plain at equal error; % MFMA peak"
PEAK_TFLOPS = {"fp64": 78.6, "fp32": 157.3, "bf16": 2500.0}
PEAK_HBM_GBS = 8000.0  # MI355X HBM3E


def peak_name(prec: str) -> str:
    """MFMA peak that bounds the trailing update of a precision mode (the
    three-product split-BF16 mode is priced against the BF16 peak with its
    algorithmic flops counted once)."""
    return "fp64" if prec == "fp64" else ("bf16" if prec.startswith("bf16") else "fp32")


def ls_flops(m: int, n: int) -> float:
    return 2.0 * m * n * n - 2.0 * n ** 3 / 3.0 + 4.0 * m * n - 1.0 * n * n


def haar(M, h, rows: int, cols: int, seed: int):
    """Haar-distributed orthonormal columns: Q diag(sign(diag R)) of a seeded
    Gaussian, factored with this library's own QR."""
    import torch
    G = np.random.default_rng(seed).standard_normal((rows, cols))
    dA = M.colmajor(G)
    tau, T = M.geqrf(h, dA)
    Q = M.orgqr(h, dA, T, cols, nq=cols)
    d = torch.sign(torch.diagonal(dA[:cols, :cols]))
    d[d == 0] = 1.0
    Q = Q * d[None, :]
    return Q


def make_problem(M, h, m: int, n: int, seed: int = 1234, smin: float = 1e-6):
    import torch
    U = haar(M, h, m, n, seed)
    V = haar(M, h, n, n, seed + 1)
    s = torch.logspace(0.0, float(np.log10(smin)), n, device="cuda", dtype=torch.float64)
    Us = M.colmajor((U * s[None, :]).contiguous())
    A = M.empty_colmajor(m, n)
    M.gemm(h, "N", "T", 1.0, Us, M.colmajor(V.contiguous()), 0.0, A)
    rng = np.random.default_rng(seed + 2)
    x0 = torch.from_numpy(rng.standard_normal(n)).cuda()
    noise = torch.from_numpy(rng.standard_normal(m)).cuda()
    b = (A @ x0) + 1e-10 * noise
    h.synchronize()
    del Us
    return {"A": A, "b": b.contiguous(), "x0": x0, "U": U, "V": V, "s": s}


def backward_errors(P, x):
    """cheap: ||A^T r|| / (||A||_F (||A||_F ||x|| + ||b||));
    kw: Karlson-Walden estimate ||(A^T A + phi^2 I)^{-1/2} A^T r|| / ||x|| / ||A||_F,
    exact through the known SVD A = U diag(s) V^T."""
    import torch
    A, b, U, s = P["A"], P["b"], P["U"], P["s"]
    r = b - A @ x
    Atr = A.T @ r
    nA = torch.linalg.norm(A)
    nx, nr, nb = torch.linalg.norm(x), torch.linalg.norm(r), torch.linalg.norm(b)
    cheap = torch.linalg.norm(Atr) / (nA * (nA * nx + nb))
    phi = nr / nx
    y = s * (U.T @ r)
    kw = torch.linalg.norm(y / torch.sqrt(s * s + phi * phi)) / nx / nA
    return {"cheap": float(cheap), "kw": float(kw), "resid": float(nr)}


# The kernel the roofline object describes: the trailing-update GEMM
# (C -= Y W), which carries half of the factorisation's flops and, with
# W = V^T C, all of its MFMA work.  Its launches are timed with HIP events on
# their own stream inside the timed region; no other kernel is, so the
# stream's dispatch gaps stay as they are.  The bulk launches count: the
# narrow ones that go first (the next panel's 64 columns, or the next outer
# block's 256 in the FP64 far update) are timed as look-ahead.
ROOF_KIND = "update_gemm"
# HBM bytes moved per algorithmic byte by the FP64 trailing update (the
# K = 256 update of each 256-column outer block), from two rocprofv3 --pmc
# runs (FETCH_SIZE, WRITE_SIZE) scaled by a 537 MB device copy in the same
# run (profiles/mfma_hbm_counters_ls16384x4096.md).
TRAFFIC_RATIO_FP64_UPDATE = 1.575
# The same for the FP32 trailing update of the mixed solve (K = 64).
TRAFFIC_RATIO_FP32_UPDATE = 1.125


def run_plain(M, h, P, steps: int, warmup: int):
    """Time `steps` plain FP64 solves, then profile one more (all kinds)."""
    import torch
    A0, b0 = P["A"], P["b"]
    m, n = A0.shape
    Aw = M.empty_colmajor(m, n)
    bw = M.empty_colmajor(m, 1)

    def step():
        Aw.copy_(A0)
        bw[:, 0].copy_(b0)
        M.gels(h, Aw, bw)

    for _ in range(warmup):
        step()
    torch.cuda.synchronize()
    M.profile_enable(h, True, kinds=[ROOF_KIND])
    t0 = time.perf_counter()
    for _ in range(steps):
        step()
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / steps
    timed = M.profile_read(h)
    x = bw[:n, 0].clone()
    M.profile_enable(h, True)
    step()
    prof = M.profile_read(h)
    M.profile_enable(h, False)
    return dt, x, timed, prof


def run_mixed(M, h, P, prec: str, steps: int, warmup: int):
    """Time `steps` mixed-precision solves, then profile one more."""
    import torch
    A0, b0 = P["A"], P["b"]
    infos = []
    for _ in range(warmup):
        M.gels_mixed(h, prec, A0, b0)
    torch.cuda.synchronize()
    M.profile_enable(h, True, kinds=[ROOF_KIND])
    t0 = time.perf_counter()
    for _ in range(steps):
        x, info = M.gels_mixed(h, prec, A0, b0)
        infos.append(info)
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / steps
    timed = M.profile_read(h)
    M.profile_enable(h, True)
    M.gels_mixed(h, prec, A0, b0)
    prof = M.profile_read(h)
    M.profile_enable(h, False)
    last = infos[-1]
    return dt, x, timed, prof, {"iterations": last.iterations, "converged": last.converged,
                                "fell_back": last.fell_back,
                                "backward_error_est": last.backward_error}


def roofline_from(timed: dict, prof: dict, prec: str) -> dict:
    """Roofline of the trailing-update GEMM from its launches in the timed
    region (algorithmic bytes and flops per launch / mean HIP-event
    duration), plus the per-kind breakdown of one separately profiled solve.
    The bound is the one the tile's algorithmic intensity selects: HBM when
    flops per byte lie below the ridge point peak_flops / 8 TB/s.  The FP32
    update at K = 64 does (16 flop/B against 19.7); the FP64 update at
    K = 256 does not (30 flop/B against 9.8) and is priced against the FP64
    MFMA peak."""
    p = timed.get(ROOF_KIND)
    if not p:
        return None
    launches = max(p["launches"], 1)
    avg_ms = p["total_ms"] / launches
    flops_per_launch = p["flops"] / launches
    bytes_per_launch = p["bytes"] / launches
    tflops = flops_per_launch / (avg_ms * 1e-3) / 1e12
    gbs = bytes_per_launch / (avg_ms * 1e-3) / 1e9
    peak_flops = PEAK_TFLOPS[peak_name(prec)]
    intensity = flops_per_launch / max(bytes_per_launch, 1.0)
    ridge = peak_flops * 1e12 / (PEAK_HBM_GBS * 1e9)
    hbm = intensity < ridge
    ratio = {"fp64": TRAFFIC_RATIO_FP64_UPDATE, "fp32": TRAFFIC_RATIO_FP32_UPDATE}.get(prec)
    traffic = round(bytes_per_launch * ratio) if ratio else None
    return {"kernel": ROOF_KIND, "bound": "hbm" if hbm else "mfma",
            "achieved": round(gbs, 1) if hbm else round(tflops, 3),
            "peak": PEAK_HBM_GBS if hbm else peak_flops,
            "unit": "GB/s" if hbm else "TFLOP/s",
            "frac": round(gbs / PEAK_HBM_GBS if hbm else tflops / peak_flops, 4),
            "traffic": traffic,
            "traffic_source": (f"measured HBM bytes / algorithmic bytes of the {prec.upper()} update "
                               f"tile ({ratio}) x this run's algorithmic bytes" if traffic else None),
            "hbm_gbs_measured_traffic": round(traffic / (avg_ms * 1e-3) / 1e9, 1) if traffic else None,
            "intensity_flop_per_byte": round(intensity, 3), "ridge_flop_per_byte": round(ridge, 3),
            "achieved_tflops": round(tflops, 3), "frac_mfma_peak": round(tflops / peak_flops, 4),
            "avg_launch_ms": round(avg_ms, 5), "launches_timed": p["launches"],
            "flops_per_launch": flops_per_launch,
            "bytes_per_launch": bytes_per_launch,
            "kernels_one_solve": {k: {"ms": round(v["total_ms"], 4), "launches": v["launches"],
                                      "tflops": round(v["flops"] / max(v["total_ms"], 1e-9) / 1e9, 3)}
                                  for k, v in prof.items()}}


def host_cpu() -> dict:
    """The host's CPU model and logical CPU count (from /proc/cpuinfo and
    os.cpu_count()), stated beside every CPU baseline."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"model": model, "logical_cpus": os.cpu_count()}


def cpu_baseline(m: int, n: int) -> dict:
    """The oracle's blocked Householder least squares at the bench's own
    m x n, one thread, one run: the reference's AVX-512 (or AVX2) blocking
    with the packed micro-kernel's accumulators in vector registers (the
    oracle's SIMD build, bit-identical to its scalar restatement)."""
    from oracle.oracle import Oracle, cpu_has_avx512
    o = Oracle("auto_simd")
    rng = np.random.default_rng(7)
    A = np.asfortranarray(rng.standard_normal((m, n)))
    b = rng.standard_normal(m)
    t0 = time.perf_counter()
    o.gels(A, b)
    t = time.perf_counter() - t0
    blocking = ("AVX-512 blocking (BS 48, 16x12 register tile)" if o.variant.startswith("avx512")
                else "AVX2 blocking (BS 96, 8x6 register tile)")
    return {"value": round(ls_flops(m, n) / t / 1e12, 5), "unit": "TFLOP/s", "cores": 1,
            "kind": "port", "seconds": round(t, 2),
            "sample": f"oracle blocked Householder LS {m}x{n} (the full workload size), iid N(0,1), "
                      f"FP64, 1 thread, {blocking}, one run",
            "cpu_avx512": bool(cpu_has_avx512()), "host": host_cpu()}


def gemm_rates(M, h) -> list:
    """FP64 MFMA GEMM tile (mqr_dgemm) at the factorisation's shapes, timed
    with events on the handle's stream (torch's current stream)."""
    import torch
    out = []
    shapes = [("square", "N", "N", 4096, 4096, 4096), ("update_k64", "N", "N", 16384, 4096, 64),
              ("update_k256", "N", "N", 16384, 4096, 256), ("vtc_k16384", "T", "N", 256, 4096, 16384)]
    for name, ta, tb, m, n, k in shapes:
        A = M.colmajor(torch.randn(k, m, dtype=torch.float64, device="cuda") if ta == "T"
                       else torch.randn(m, k, dtype=torch.float64, device="cuda"))
        B = M.colmajor(torch.randn(k, n, dtype=torch.float64, device="cuda"))
        Cm = M.colmajor(torch.zeros(m, n, dtype=torch.float64, device="cuda"))
        for _ in range(2):
            M.gemm(h, ta, tb, 1.0, A, B, 1.0, Cm)
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        reps = 5
        e0.record()
        for _ in range(reps):
            M.gemm(h, ta, tb, 1.0, A, B, 1.0, Cm)
        e1.record()
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1) / reps
        tf = 2.0 * m * n * k / (ms * 1e-3) / 1e12
        out.append({"shape": name, "m": m, "n": n, "k": k, "ms": round(ms, 4), "tflops": round(tf, 2),
                    "frac_fp64_peak": round(tf / PEAK_TFLOPS["fp64"], 3)})
        print(f"[bench] dgemm {name} {m}x{n}x{k}: {ms:.3f} ms {tf:.2f} TFLOP/s", file=sys.stderr,
              flush=True)
    return out


def qr_flops(m: int, n: int) -> float:
    """dgeqrf count 2mn^2 - 2n^3/3 (m >= n)."""
    return 2.0 * m * n * n - 2.0 * n ** 3 / 3.0


def orgqr_flops(m: int, n: int, k: int) -> float:
    """Thin Q (m x n from k reflectors): 4mnk - 2(m+n)k^2 + 4k^3/3."""
    return 4.0 * m * n * k - 2.0 * (m + n) * k * k + 4.0 * k ** 3 / 3.0


def qr_rates(M, h, steps: int = 3, seed: int = 2024) -> list:
    """Plain FP64 QR (mqr_dgeqrf) of square and tall matrices, iid N(0,1) and
    graded A = U diag(logspace(0, -8)) V^T (Haar U, V); the factorisation and
    the thin Q (mqr_dorgqr) are timed separately with events on the handle's
    stream (torch's current stream), and checked by ||QR - A||_F / ||A||_F and
    ||Q^T Q - I||_F."""
    import torch
    out = []
    cases = [("iid", 8192, 8192, None), ("graded", 8192, 8192, 1e-8),
             ("iid", 32768, 1024, None), ("graded", 32768, 1024, 1e-8)]
    for kind, m, n, smin in cases:
        if smin is None:
            A0 = M.colmajor(torch.from_numpy(
                np.random.default_rng(seed).standard_normal((m, n))).cuda())
        else:
            U = haar(M, h, m, n, seed)
            V = haar(M, h, n, n, seed + 1)
            sv = torch.logspace(0.0, float(np.log10(smin)), n, device="cuda",
                                dtype=torch.float64)
            Us = M.colmajor((U * sv[None, :]).contiguous())
            A0 = M.empty_colmajor(m, n)
            M.gemm(h, "N", "T", 1.0, Us, M.colmajor(V.contiguous()), 0.0, A0)
            del U, V, Us
        k = min(m, n)
        Aw = M.empty_colmajor(m, n)
        f_ms, q_ms = [], []
        for rep in range(steps + 1):
            Aw.copy_(A0)
            torch.cuda.synchronize()
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record()
            tau, T = M.geqrf(h, Aw)
            e1.record()
            torch.cuda.synchronize()
            if rep:
                f_ms.append(e0.elapsed_time(e1))
        for rep in range(steps + 1):
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record()
            Q = M.orgqr(h, Aw, T, k, nq=k)
            e1.record()
            torch.cuda.synchronize()
            if rep:
                q_ms.append(e0.elapsed_time(e1))
            if rep < steps:
                del Q
        R = torch.triu(Aw[:k, :])
        res = float(torch.linalg.norm(Q @ R - A0) / torch.linalg.norm(A0))
        orth = float(torch.linalg.norm(Q.T @ Q - torch.eye(k, device="cuda", dtype=torch.float64)))
        tf = float(np.median(f_ms))
        tq = float(np.median(q_ms))
        rec = {"case": f"{m}x{n} {kind}", "m": m, "n": n, "seed": seed,
               "factor_ms": round(tf, 3),
               "factor_tflops": round(qr_flops(m, n) / (tf * 1e-3) / 1e12, 3),
               "frac_fp64_peak": round(qr_flops(m, n) / (tf * 1e-3) / 1e12 / PEAK_TFLOPS["fp64"], 4),
               "thin_q_ms": round(tq, 3),
               "thin_q_tflops": round(orgqr_flops(m, k, k) / (tq * 1e-3) / 1e12, 3),
               "qr_residual": res, "orthogonality": orth}
        print(f"[bench] qr {rec['case']}: factor {tf:.3f} ms {rec['factor_tflops']} TFLOP/s, "
              f"thin Q {tq:.3f} ms {rec['thin_q_tflops']} TFLOP/s, |QR-A|/|A| {res:.2e}, "
              f"|QtQ-I| {orth:.2e}", file=sys.stderr, flush=True)
        out.append(rec)
        del A0, Aw, Q, R, tau, T
        torch.cuda.empty_cache()
    return out


def cpu_qr_baseline(n: int = 1024, runs: int = 5) -> dict:
    """The oracle's QR(A, Q, R, QR_Householder) at n x n iid N(0,1): blocked
    factorisation and the full Q, one thread, median of `runs`."""
    from oracle.oracle import Oracle
    o = Oracle("auto_simd")
    A = np.asfortranarray(np.random.default_rng(11).standard_normal((n, n)))
    times = []
    for _ in range(runs):
        t0 = time.perf_counter()
        o.qr(A)
        times.append(time.perf_counter() - t0)
    t = float(np.median(times))
    return {"value": round(qr_flops(n, n) / t / 1e12, 5), "unit": "TFLOP/s (factor flops)",
            "cores": 1, "kind": "port",
            "sample": f"oracle QR(A, Q, R) {n}x{n} iid N(0,1), factor + full Q, FP64, 1 thread, "
                      f"{o.variant} blocking, median of {runs} runs ({t:.3f} s each)",
            "host": host_cpu()}


def batched_shapes(count: int = 10000, seed: int = 4321):
    """(m, n) of the batched workload: m ~ U{16..512}, n ~ U{8..m}, seeded."""
    rng = np.random.default_rng(seed)
    ms = rng.integers(16, 513, count)
    ns = np.array([rng.integers(8, m + 1) for m in ms])
    return ms, ns


def _cpu_factor_share(job) -> float:
    """Seconds the oracle's blocked factorisation takes over the given shapes
    (iid N(0,1) matrices generated before the clock starts)."""
    shapes, seed = job
    from oracle.oracle import Oracle
    o = Oracle("auto_simd")
    rng = np.random.default_rng(seed)
    mats = [np.asfortranarray(rng.standard_normal((m, n))) for m, n in shapes]
    t0 = time.perf_counter()
    for A in mats:
        o.geqrf_blocked(A)
    return time.perf_counter() - t0


def cpu_batched_baseline(sample: int = 1000) -> dict:
    """The oracle's blocked factorisation, one call per matrix (the reference's
    QR loop), over the first `sample` matrices of the batched workload: one
    thread, and one process per core of the box's share (at most 16) with
    the sample dealt out round-robin, whose slowest process sets the time.
    Call it before anything touches the GPU: the pool starts new processes."""
    import multiprocessing as mp
    ms, ns = batched_shapes()
    shapes = [(int(m), int(n)) for m, n in zip(ms[:sample], ns[:sample])]
    flops = float(sum(2.0 * m * n * n - 2.0 * n ** 3 / 3.0 for m, n in shapes))
    t1 = _cpu_factor_share((shapes, 1))
    procs = max(1, min(16, os.cpu_count() or 1))
    with mp.get_context("spawn").Pool(procs) as pool:
        tp = max(pool.map(_cpu_factor_share, [(shapes[i::procs], 2 + i) for i in range(procs)]))
    sample_desc = (f"oracle blocked factorisation (geqrf) of the first {sample} matrices of the "
                   f"batch (same seeded shapes, iid N(0,1)), FP64")
    return {"single_thread": {"value": round(flops / t1 / 1e12, 5), "unit": "TFLOP/s",
                              "cores": 1, "kind": "port", "seconds": round(t1, 3),
                              "sample": sample_desc + ", 1 thread"},
            "per_core": {"value": round(flops / tp / 1e12, 5), "unit": "TFLOP/s",
                         "cores": procs, "kind": "port", "seconds": round(tp, 3),
                         "sample": sample_desc + f", {procs} processes"},
            "host": host_cpu()}


def cpu_qr_large_baseline(n: int = 8192) -> dict:
    """The oracle's blocked factorisation of one n x n iid N(0,1) matrix
    (the 8192^2 configuration), one thread, timed once."""
    from oracle.oracle import Oracle
    o = Oracle("auto_simd")
    A = np.asfortranarray(np.random.default_rng(12).standard_normal((n, n)))
    t0 = time.perf_counter()
    o.geqrf_blocked(A)
    t = time.perf_counter() - t0
    return {"value": round(qr_flops(n, n) / t / 1e12, 5), "unit": "TFLOP/s (factor flops)",
            "cores": 1, "kind": "port", "seconds": round(t, 2),
            "sample": f"oracle blocked factorisation {n}x{n} iid N(0,1), FP64, 1 thread, "
                      f"{o.variant} blocking, one run", "host": host_cpu()}


def batched_rates(M, h, count: int = 10000, seed: int = 4321) -> list:
    """Batched uneven QR (m ~ U{16..512}, n ~ U{8..m}, iid N(0,1)) in FP64 and
    FP32: one persistent launch replayed from its hipGraph, timed with events
    on the handle's stream around the replay (inputs restored before each)."""
    import numpy as np
    import torch
    ms, ns = batched_shapes(count, seed)
    flops = float(np.sum(2.0 * ms * ns * ns - 2.0 * ns ** 3 / 3.0))
    out = []
    g = torch.Generator(device="cuda").manual_seed(seed)
    for dt, name in ((torch.float64, "fp64"), (torch.float32, "fp32")):
        src = [torch.randn(int(n), int(m), generator=g, device="cuda", dtype=dt).T
               for m, n in zip(ms, ns)]
        work = [torch.empty(int(n), int(m), device="cuda", dtype=dt).T for m, n in zip(ms, ns)]
        plan = M.BatchedPlan(h, work)
        times = []
        for rep in range(4):
            for w, s in zip(work, src):
                w.copy_(s)
            torch.cuda.synchronize()
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record()
            plan.run()
            e1.record()
            torch.cuda.synchronize()
            times.append(e0.elapsed_time(e1))
        plan.close()
        t = float(np.median(times[1:]))
        peak = PEAK_TFLOPS[name]
        rec = {"dtype": name, "count": count, "ms": round(t, 3),
               "tflops": round(flops / (t * 1e-3) / 1e12, 3),
               "frac_mfma_peak": round(flops / (t * 1e-3) / 1e12 / peak, 4), "flops": flops}
        print(f"[bench] batched {name}: {t:.3f} ms {rec['tflops']} TFLOP/s", file=sys.stderr,
              flush=True)
        out.append(rec)
        del src, work
    return out


def breakdown_rates(M, h, m: int = 16384, n: int = 4096, steps: int = 3, seed: int = 5) -> list:
    """mqr_dgeqrf of m x n inputs whose panels defeat CholeskyQR: iid
    N(0,1) as the reference point; every odd column a copy of the even one
    before it (each panel breaks down and is factored column by column by
    the row-chunk kernel's workgroups); columns scaled by logspace(0, -14)
    (ill-conditioned, yet each panel stays well conditioned after the
    earlier ones are removed).  Events on the handle's stream; the best of
    `steps` runs after one warm-up, with ||QR - A|| / ||A|| of the last."""
    import torch
    g = torch.Generator(device="cuda").manual_seed(seed)
    base = torch.randn(m, n, generator=g, device="cuda", dtype=torch.float64)
    dup = base.clone()
    dup[:, 1::2] = dup[:, 0::2]
    scaled = base * torch.logspace(0, -14, n, device="cuda", dtype=torch.float64)[None, :]
    out = []
    for name, A in (("iid", base), ("duplicated_columns", dup), ("columns_scaled_1e-14", scaled)):
        A0 = M.colmajor(A)
        Aw = M.empty_colmajor(m, n)
        best = float("inf")
        for rep in range(steps + 1):
            Aw.copy_(A0)
            torch.cuda.synchronize()
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record()
            tau, T = M.geqrf(h, Aw)
            e1.record()
            torch.cuda.synchronize()
            if rep > 0:
                best = min(best, e0.elapsed_time(e1))
        Q = M.orgqr(h, Aw, T, n, nq=n)
        h.synchronize()
        res = float(torch.linalg.norm(Q @ torch.triu(Aw[:n, :]) - A0) / torch.linalg.norm(A0))
        rec = {"input": name, "m": m, "n": n, "factor_ms": round(best, 3),
               "tflops": round(qr_flops(m, n) / (best * 1e-3) / 1e12, 3),
               "rel_residual": res}
        print(f"[bench] breakdown {name}: {best:.3f} ms |QR-A|/|A| {res:.2e}", file=sys.stderr,
              flush=True)
        out.append(rec)
        del A0, Aw, Q, tau, T
    return out


def adjoint_rates(M, h, batch: int = 256, m: int = 1024, n: int = 256, seed: int = 99) -> list:
    """QR adjoint (mqr_qr_backward) on a batch of thin QRs of iid N(0,1)
    matrices, FP32 and FP64, loss sum(Wq o Q) + sum(Wr o R) so that dQ = Wq,
    dR = Wr; events on the handle's stream around the backward call only."""
    import torch
    out = []
    g = torch.Generator(device="cuda").manual_seed(seed)
    A = torch.randn(batch, m, n, generator=g, device="cuda", dtype=torch.float64)
    Q, R = torch.linalg.qr(A.cpu(), mode="reduced")
    flops = batch * (5.0 * m * n * n + 2.0 * n ** 3 / 3.0)
    for dt, name in ((torch.float32, "fp32"), (torch.float64, "fp64")):
        Qd = M.colmajor_batched(Q.to(device="cuda", dtype=dt))
        Rd = M.colmajor_batched(R.to(device="cuda", dtype=dt))
        Wq = M.colmajor_batched(torch.randn(batch, m, n, generator=g, device="cuda", dtype=dt))
        Wr = M.colmajor_batched(torch.randn(batch, n, n, generator=g, device="cuda", dtype=dt))
        for _ in range(2):
            M.qr_backward(h, Qd, Rd, Wq, Wr)
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        reps = 5
        e0.record()
        for _ in range(reps):
            M.qr_backward(h, Qd, Rd, Wq, Wr)
        e1.record()
        torch.cuda.synchronize()
        t = e0.elapsed_time(e1) / reps
        rec = {"dtype": name, "batch": batch, "m": m, "n": n, "ms": round(t, 3),
               "tflops": round(flops / (t * 1e-3) / 1e12, 3),
               "frac_mfma_peak": round(flops / (t * 1e-3) / 1e12 / PEAK_TFLOPS[name], 4)}
        print(f"[bench] adjoint {name}: {t:.3f} ms {rec['tflops']} TFLOP/s", file=sys.stderr,
              flush=True)
        out.append(rec)
    return out


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--m", type=int, default=16384)
    ap.add_argument("--n", type=int, default=4096)
    ap.add_argument("--nb", type=int, default=None)
    ap.add_argument("--prec", default="fp32", choices=["fp32", "bf16x3", "bf16", "none"],
                    help="mixed-precision trailing updates ('none': plain FP64 only)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--kernel", choices=["gemm", "batched", "adjoint", "qr", "breakdown"],
                    default=None, help="time one kernel family instead of the solve")
    args = ap.parse_args()
    if args.kernel == "breakdown":
        import mat_amd as M
        h = M.Handle(0)
        print(json.dumps({"qr_breakdown": breakdown_rates(M, h)}))
        return 0
    if args.kernel == "adjoint":
        import mat_amd as M
        h = M.Handle(0)
        print(json.dumps({"qr_adjoint": adjoint_rates(M, h)}))
        return 0
    if args.kernel == "batched":
        # the CPU pool first: nothing has touched the GPU yet
        cpu = None if args.no_cpu else cpu_batched_baseline()
        import mat_amd as M
        h = M.Handle(0)
        rec = {"batched_qr": batched_rates(M, h)}
        if cpu is not None:
            rec["cpu_baseline"] = cpu
        print(json.dumps(rec))
        return 0
    if args.kernel == "qr":
        import mat_amd as M
        h = M.Handle(0)
        rec = {"qr_fp64": qr_rates(M, h)}
        if not args.no_cpu:
            rec["cpu_baseline"] = cpu_qr_baseline()
            rec["cpu_baseline_large"] = cpu_qr_large_baseline()
        print(json.dumps(rec))
        return 0
    if args.kernel == "gemm":
        import mat_amd as M
        h = M.Handle(0)
        print(json.dumps({"dgemm": gemm_rates(M, h)}))
        return 0
    if args.gpus != 1:
        print(f"bench.py measures one GPU only: --gpus {args.gpus} is not supported "
              "(the path runs on a single MI355X; see the sharding design notes)", file=sys.stderr)
        return 1

    import torch
    import mat_amd as M
    assert torch.cuda.is_available(), "no GPU"
    h = M.Handle(0, nb=args.nb)
    m, n = args.m, args.n
    print(f"[bench] generating {m}x{n} graded problem", file=sys.stderr, flush=True)
    P = make_problem(M, h, m, n)
    flops = ls_flops(m, n)

    print("[bench] plain FP64", file=sys.stderr, flush=True)
    dt_p, x_p, timed_p, prof_p = run_plain(M, h, P, args.steps, args.warmup)
    be_p = backward_errors(P, x_p)
    plain = {"tflops": round(flops / dt_p / 1e12, 4), "ms": round(dt_p * 1e3, 3),
             "pct_fp64_peak": round(100 * flops / dt_p / 1e12 / PEAK_TFLOPS["fp64"], 2),
             "backward_error": be_p, "roofline": roofline_from(timed_p, prof_p, "fp64")}
    print(f"[bench] plain: {plain['tflops']} TFLOP/s, {plain['ms']} ms, be={be_p}",
          file=sys.stderr, flush=True)

    mixed = None
    if args.prec != "none":
        print(f"[bench] mixed {args.prec} + refinement", file=sys.stderr, flush=True)
        try:
            dt_m, x_m, timed_m, prof_m, ir = run_mixed(M, h, P, args.prec, args.steps,
                                                       args.warmup)
            be_m = backward_errors(P, x_m)
            mixed = {"prec": args.prec, "tflops": round(flops / dt_m / 1e12, 4),
                     "ms": round(dt_m * 1e3, 3), "backward_error": be_m, "ir": ir,
                     "roofline": roofline_from(timed_m, prof_m, args.prec)}
            print(f"[bench] mixed: {mixed['tflops']} TFLOP/s, {mixed['ms']} ms, be={be_m}, ir={ir}",
                  file=sys.stderr, flush=True)
        except M.MqrError as e:
            print(f"[bench] mixed path unavailable: {e}", file=sys.stderr, flush=True)

    head = mixed if mixed is not None else plain
    dt_head = head["ms"] / 1e3
    cpu = None if args.no_cpu else cpu_baseline(m, n)
    rec = {
        "metric": METRIC,
        "value": head["tflops"],
        "unit": "TFLOP/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": head["ms"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if mixed is None else f"{args.prec}+f64",
        "data": "synthetic",
        "config": {"workload": f"least squares {m}x{n}, A = U diag(logspace(0,-6)) V^T (Haar U, V), "
                               f"b = A x0 + 1e-10 noise; one full solve per step",
                   "m": m, "n": n, "nb": h.nb, "mixed_prec": None if mixed is None else args.prec,
                   "seed": 1234},
        "pct_mfma_peak": round(100 * head["tflops"] / PEAK_TFLOPS[peak_name("fp64" if mixed is None else args.prec)], 2),
        "speedup_vs_fp64_plain": None if mixed is None else round(plain["ms"] / mixed["ms"], 3),
        "fp64_plain": plain,
        "mixed": mixed,
        "roofline": head["roofline"],
        "cpu_baseline": cpu,
    }
    print(json.dumps(rec))
    h.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
