"""Autotuner of the batched uneven QR's tile (panel width 8 or 16) per size
bin and precision.

    python -m mat_amd.autotune_batched --out mat_amd/csrc/tuned_batched.h

For every bin of m (16-64, 65-128, 129-256, 257-512) and both precisions it
builds a batch of `--count` matrices (m uniform in the bin, n uniform in
[8, m], iid N(0,1)), captures the persistent launch once per forced tile
(mqr_batched_plan_create with tile 0 or 1), times `--reps` replays with HIP
events and writes the faster width into a C header that batched.hip
compiles in (the library's tile = -1 path).  Needs a GPU; rebuild the
library afterwards.
"""
from __future__ import annotations

import argparse
import json
import sys

BINS = [(16, 64), (65, 128), (129, 256), (257, 512)]
WIDTHS = {0: 8, 1: 16}


def time_plan(M, h, mats, tile: int, reps: int) -> float:
    import torch
    src = [t.clone() for t in mats]
    plan = M.BatchedPlan(h, mats, tile=tile)
    times = []
    for _ in range(reps + 1):
        for d, s in zip(mats, src):
            d.copy_(s)
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        plan.run()
        e1.record()
        torch.cuda.synchronize()
        times.append(e0.elapsed_time(e1))
    plan.close()
    times = sorted(times[1:])
    return times[len(times) // 2]


def tune(count: int, reps: int, seed: int) -> dict:
    import numpy as np
    import torch
    import mat_amd as M
    h = M.Handle(0)
    g = torch.Generator(device="cuda").manual_seed(seed)
    rng = np.random.default_rng(seed)
    result = {}
    for dt, name in ((torch.float32, "float"), (torch.float64, "double")):
        rows = []
        for lo, hi in BINS:
            ms = rng.integers(lo, hi + 1, count)
            mats = [torch.randn(int(rng.integers(8, m + 1)), int(m), generator=g, device="cuda",
                                dtype=dt).T for m in ms]
            t = {tile: time_plan(M, h, mats, tile, reps) for tile in WIDTHS}
            best = min(t, key=t.get)
            rows.append({"bin": [lo, hi], "ms": {WIDTHS[k]: round(v, 4) for k, v in t.items()},
                         "nb": WIDTHS[best]})
            print(f"[autotune] {name} m in [{lo}, {hi}]: " +
                  ", ".join(f"nb {WIDTHS[k]}: {v:.3f} ms" for k, v in t.items()) +
                  f" -> {WIDTHS[best]}", file=sys.stderr, flush=True)
        result[name] = rows
    h.close()
    return result


def header(result: dict, count: int, reps: int) -> str:
    lines = ["// Generated by mat_amd/autotune_batched.py on one MI355X: panel width of",
             f"// the batched uneven QR per m bin, the faster of 8 and 16 on {count} matrices",
             f"// per bin (n uniform in [8, m]), median of {reps} hipGraph replays.", "//"]
    for name in ("float", "double"):
        for r in result[name]:
            lines.append(f"//   {name:6s} m in [{r['bin'][0]:3d}, {r['bin'][1]:3d}]: " +
                         ", ".join(f"nb {k} {v} ms" for k, v in r["ms"].items()))
    lines += ["#pragma once", "",
              "static const int kTunedNb[2][4] = {",
              "    /* float  */ {" + ", ".join(str(r["nb"]) for r in result["float"]) + "},",
              "    /* double */ {" + ", ".join(str(r["nb"]) for r in result["double"]) + "},",
              "};", ""]
    return "\n".join(lines)


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default=None, help="header to write (default: print only)")
    ap.add_argument("--count", type=int, default=1500)
    ap.add_argument("--reps", type=int, default=3)
    ap.add_argument("--seed", type=int, default=77)
    args = ap.parse_args()
    result = tune(args.count, args.reps, args.seed)
    text = header(result, args.count, args.reps)
    if args.out:
        with open(args.out, "w") as f:
            f.write(text)
    print(json.dumps(result))
    return 0


if __name__ == "__main__":
    sys.exit(main())
