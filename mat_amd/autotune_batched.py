"""Autotuner of the batched uneven QR's launch (panel width per size bin,
residency and queue order) per precision.

    python -m mat_amd.autotune_batched --out mat_amd/csrc/tuned_batched.h

Two stages, both timed as hipGraph replays (mqr_batched_plan_create_sched)
with HIP events, median of `--reps` after a warm-up, on fresh copies of the
inputs:

1. Panel width.  For every bin of m (16-64, 65-128, 129-256, 257-512) a
   batch of `--count` matrices (m uniform in the bin, n uniform in [8, m],
   iid N(0,1)) is factored with widths 8 and 16; the faster one is the bin's.
2. Schedule.  On a mixed batch of `--mixed` matrices (m ~ U{16..512},
   n ~ U{8..m}, the shape of the batched workload) with the widths of stage
   1, every pair of residency (1 .. the kernel's occupancy limit workgroups
   per CU) and queue order (decreasing flop count; as given; largest and
   smallest alternating) is timed; the fastest pair is kept.

The result goes into a C header that batched.hip compiles in (the
library's tuned path, tile = -1 / schedule fields -1).  Needs a GPU;
rebuild the library afterwards.
"""
from __future__ import annotations

import argparse
import json
import sys

BINS = [(16, 64), (65, 128), (129, 256), (257, 512)]
WIDTHS = {0: 8, 1: 16}
ORDERS = {0: "longest first", 1: "as given", 2: "largest/smallest alternating"}


def time_plan(M, h, mats, reps: int, tile: int = -1, per_cu=None, order=None) -> float:
    import torch
    src = [t.clone() for t in mats]
    plan = M.BatchedPlan(h, mats, tile=tile, per_cu=per_cu, order=order)
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
    for d, s in zip(mats, src):
        d.copy_(s)
    times = sorted(times[1:])
    return times[len(times) // 2]


def make_batch(M, rng, g, dt, count: int, lo: int, hi: int):
    import torch
    ms = rng.integers(lo, hi + 1, count)
    return [torch.randn(int(rng.integers(8, m + 1)), int(m), generator=g, device="cuda",
                        dtype=dt).T for m in ms]


def tune(count: int, mixed: int, reps: int, seed: int) -> dict:
    import numpy as np
    import torch
    import mat_amd as M
    h = M.Handle(0)
    g = torch.Generator(device="cuda").manual_seed(seed)
    rng = np.random.default_rng(seed)
    result = {}
    for dt, name, max_cu in ((torch.float32, "float", 3), (torch.float64, "double", 2)):
        rows = []
        for lo, hi in BINS:
            mats = make_batch(M, rng, g, dt, count, lo, hi)
            t = {tile: time_plan(M, h, mats, reps, tile=tile, per_cu=0, order=0)
                 for tile in WIDTHS}
            best = min(t, key=t.get)
            rows.append({"bin": [lo, hi], "ms": {WIDTHS[k]: round(v, 4) for k, v in t.items()},
                         "nb": WIDTHS[best]})
            print(f"[autotune] {name} m in [{lo}, {hi}]: " +
                  ", ".join(f"nb {WIDTHS[k]}: {v:.3f} ms" for k, v in t.items()) +
                  f" -> {WIDTHS[best]}", file=sys.stderr, flush=True)
            del mats
        # stage 2 on the mixed batch; the bins' widths apply only through
        # the tuned path, so force the width the bins chose when they agree
        widths = {r["nb"] for r in rows}
        tile = {8: 0, 16: 1}[widths.pop()] if len(widths) == 1 else -1
        mats = make_batch(M, rng, g, dt, mixed, 16, 512)
        sched = {}
        for per_cu in range(1, max_cu + 1):
            for order in ORDERS:
                ms = time_plan(M, h, mats, reps, tile=tile, per_cu=per_cu, order=order)
                sched[(per_cu, order)] = ms
                print(f"[autotune] {name} mixed {mixed}: {per_cu} per CU, {ORDERS[order]}: "
                      f"{ms:.3f} ms", file=sys.stderr, flush=True)
        (bc, bo) = min(sched, key=sched.get)
        del mats
        result[name] = {"bins": rows, "per_cu": bc, "order": bo,
                        "schedules": {f"{c}/{o}": round(v, 4) for (c, o), v in sched.items()}}
    h.close()
    return result


def header(result: dict, count: int, mixed: int, reps: int) -> str:
    lines = ["// Generated by mat_amd/autotune_batched.py on one MI355X.",
             f"// Panel width of the batched uneven QR per m bin: the faster of 8 and 16 on {count}",
             "// matrices per bin (n uniform in [8, m]); residency and queue order: the fastest",
             f"// pair on {mixed} matrices with m ~ U{{16..512}}, n ~ U{{8..m}}; median of {reps} hipGraph",
             "// replays each.", "//"]
    for name in ("float", "double"):
        for r in result[name]["bins"]:
            lines.append(f"//   {name:6s} m in [{r['bin'][0]:3d}, {r['bin'][1]:3d}]: " +
                         ", ".join(f"nb {k} {v} ms" for k, v in r["ms"].items()))
        lines.append(f"//   {name:6s} schedules (per CU / order: ms): " +
                     ", ".join(f"{k} {v}" for k, v in result[name]["schedules"].items()))
    lines += ["#pragma once", "",
              "static const int kTunedNb[2][4] = {",
              "    /* float  */ {" + ", ".join(str(r["nb"]) for r in result["float"]["bins"]) + "},",
              "    /* double */ {" + ", ".join(str(r["nb"]) for r in result["double"]["bins"]) + "},",
              "};", "",
              "// Schedule: resident workgroups per CU (0: the occupancy limit) and queue",
              "// order (0: decreasing flop count, 1: as given, 2: largest and smallest",
              "// alternating), [float, double].",
              f"static const int kTunedPerCu[2] = {{{result['float']['per_cu']}, "
              f"{result['double']['per_cu']}}};",
              f"static const int kTunedOrder[2] = {{{result['float']['order']}, "
              f"{result['double']['order']}}};", ""]
    return "\n".join(lines)


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default=None, help="header to write (default: print only)")
    ap.add_argument("--count", type=int, default=1500)
    ap.add_argument("--mixed", type=int, default=5000)
    ap.add_argument("--reps", type=int, default=3)
    ap.add_argument("--seed", type=int, default=77)
    args = ap.parse_args()
    result = tune(args.count, args.mixed, args.reps, args.seed)
    text = header(result, args.count, args.mixed, args.reps)
    if args.out:
        with open(args.out, "w") as f:
            f.write(text)
    print(json.dumps(result))
    return 0


if __name__ == "__main__":
    sys.exit(main())
