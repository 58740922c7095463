"""Per-kernel summary of rocprofv3 counter-collection CSVs.

Reads the `*_counter_collection.csv` files of separate `rocprofv3 --pmc`
passes over the same program (one file per pass; the dispatch IDs of a
deterministic program agree between passes, which is checked by kernel
name), keeps the dispatches in an ID range, and prints one markdown row per
kernel with
  - launches, mean duration of one launch (the counter passes serialise
    dispatches, so this is the kernel running alone),
  - MFMA work from SQ_INSTS_VALU_MFMA_MOPS_{F64,F32,BF16} (each unit is 512
    flops), its rate and its fraction of the MI355X dense peak of that type,
  - the MFMA-busy share SQ_VALU_MFMA_BUSY_CYCLES / (GRBM_GUI_ACTIVE / XCDs *
    SIMDs), rocprofv3's own MfmaUtil formula (GRBM_GUI_ACTIVE arrives summed
    over the XCDs' instances),
  - HBM bytes from FETCH_SIZE and WRITE_SIZE (KiB per dispatch), each scaled
    by what it reads for a device copy of known size in the same run, and
    their rate against 8 TB/s.
Non-transposed GEMM tiles (`k_gemm_f64<T, false, ...>` / `k_gemm_f32<T,
false, ...>`, the C -= Y W updates) are split by their effective inner
dimension, counted MFMA flops / (2 * tiles * TILE^2), so that one template
launched at K = 64 and at K = 256 shows as two rows.

Usage:
  python3 -m mat_amd.rocprof_summary --mfma m.csv --fetch f.csv --write w.csv \
      --range 2121:4139 --per 2 --copy-bytes 536870912
"""
from __future__ import annotations

import argparse
import collections
import csv
import re
import sys

PEAK_TFLOPS = {"F64": 78.6, "F32": 157.3, "BF16": 2500.0}
PEAK_HBM_GBS = 8000.0
XCDS = 8
SIMDS = 1024
MOPS = {"F64": "SQ_INSTS_VALU_MFMA_MOPS_F64", "F32": "SQ_INSTS_VALU_MFMA_MOPS_F32",
        "BF16": "SQ_INSTS_VALU_MFMA_MOPS_BF16"}


def load(path: str) -> dict:
    """dispatch id -> {name, grid, wg, ns, counters{}}"""
    out = {}
    with open(path) as f:
        for r in csv.DictReader(f):
            d = int(r["Dispatch_Id"])
            e = out.get(d)
            if e is None:
                e = out[d] = {"name": r["Kernel_Name"], "grid": int(r["Grid_Size"]),
                              "wg": int(r["Workgroup_Size"]),
                              "ns": int(r["End_Timestamp"]) - int(r["Start_Timestamp"]),
                              "c": {}}
            e["c"][r["Counter_Name"]] = e["c"].get(r["Counter_Name"], 0.0) + float(r["Counter_Value"])
    return out


def short(name: str) -> str:
    n = re.sub(r"^void ", "", name)
    n = re.sub(r"mqr::(\(anonymous namespace\)::)?", "", n)
    n = re.sub(r"\(.*$", "", n)
    return n


def tile_of(name: str) -> int | None:
    """Tile width of a non-transposed GEMM tile (C -= Y W), else None."""
    m = re.match(r"k_gemm_f(?:64|32)<(\d+), false,", short(name))
    return int(m.group(1)) if m else None


def copy_scales(fetch: dict, write: dict, copy_bytes: float) -> tuple:
    """FETCH/WRITE scale factors (bytes per KiB counted) from the device
    copies (runtime copyBuffer) that moved copy_bytes each way."""
    fs, ws = [], []
    for d, e in write.items():
        if "copyBuffer" in e["name"] and abs(e["c"].get("WRITE_SIZE", 0) * 1024 - copy_bytes) < 0.01 * copy_bytes:
            ws.append(copy_bytes / e["c"]["WRITE_SIZE"])
            if d in fetch and fetch[d]["c"].get("FETCH_SIZE"):
                fs.append(copy_bytes / fetch[d]["c"]["FETCH_SIZE"])
    if not fs or not ws:
        raise SystemExit("no device copy of the stated size found for the FETCH/WRITE scale")
    return sum(fs) / len(fs), sum(ws) / len(ws)


def summarise(mfma: dict | None, fetch: dict | None, write: dict | None, lo: int, hi: int,
              per: float, fscale: float, wscale: float) -> list:
    base = mfma or fetch or write
    rows = collections.OrderedDict()
    for d in sorted(base):
        if not (lo <= d <= hi):
            continue
        name = base[d]["name"]
        for other in (mfma, fetch, write):
            if other is not None and d in other and other[d]["name"] != name:
                raise SystemExit(f"dispatch {d} differs between passes: {name} / {other[d]['name']}")
        key = short(name)
        flops = {}
        if mfma is not None and d in mfma:
            for t, c in MOPS.items():
                flops[t] = mfma[d]["c"].get(c, 0.0) * 512.0
        t = tile_of(name)
        if t and flops:
            tiles = max(base[d]["grid"] // max(base[d]["wg"], 1), 1)
            keff = sum(flops.values()) / (2.0 * tiles * t * t)
            key += f" [K~{int(round(keff / 16.0) * 16)}]"
        r = rows.setdefault(key, {"n": 0, "ns": 0.0, "fl": collections.Counter(), "busy": 0.0,
                                  "grbm": 0.0, "rd": 0.0, "wr": 0.0, "has_mfma": False,
                                  "has_rd": False, "has_wr": False})
        r["n"] += 1
        r["ns"] += base[d]["ns"]
        if flops:
            r["has_mfma"] = True
            r["fl"].update(flops)
            r["busy"] += mfma[d]["c"].get("SQ_VALU_MFMA_BUSY_CYCLES", 0.0)
            r["grbm"] += mfma[d]["c"].get("GRBM_GUI_ACTIVE", 0.0)
        if fetch is not None and d in fetch:
            r["has_rd"] = True
            r["rd"] += fetch[d]["c"].get("FETCH_SIZE", 0.0) * fscale
        if write is not None and d in write:
            r["has_wr"] = True
            r["wr"] += write[d]["c"].get("WRITE_SIZE", 0.0) * wscale
    out = []
    for k, r in rows.items():
        n, s = r["n"], r["ns"] * 1e-9
        rec = {"kernel": k, "launches": n / per, "avg_us": r["ns"] / n * 1e-3,
               "total_ms": r["ns"] / per * 1e-6}
        if r["has_mfma"]:
            typ = max(r["fl"], key=lambda t: r["fl"][t]) if any(r["fl"].values()) else None
            tot = sum(r["fl"].values())
            rec["mfma_type"] = typ
            rec["mfma_gflop_per_launch"] = tot / n * 1e-9
            rec["mfma_tflops"] = tot / s * 1e-12 if s > 0 else 0.0
            rec["frac_peak"] = rec["mfma_tflops"] / PEAK_TFLOPS[typ] if typ else 0.0
            rec["mfma_util_pct"] = (100.0 * r["busy"] / (r["grbm"] / XCDS * SIMDS)
                                    if r["grbm"] else 0.0)
        if r["has_rd"] or r["has_wr"]:
            rec["read_mb"] = r["rd"] / n * 1e-6
            rec["write_mb"] = r["wr"] / n * 1e-6
            rec["hbm_gbs"] = (r["rd"] + r["wr"]) / s * 1e-9 if s > 0 else 0.0
            rec["frac_hbm"] = rec["hbm_gbs"] / PEAK_HBM_GBS
        out.append(rec)
    out.sort(key=lambda x: -x["total_ms"])
    return out


def markdown(recs: list, top: int) -> str:
    head = ("| kernel | launches | avg µs (alone) | ms total | MFMA type | GFLOP/launch | TFLOP/s | "
            "frac of peak | MfmaUtil % | read MB | write MB | HBM GB/s | frac of 8 TB/s |")
    lines = [head, "|" + "---|" * 13]
    for r in recs[:top]:
        f = lambda k, fmt: (fmt % r[k]) if k in r and r[k] is not None else ""
        lines.append("| `%s` | %s | %.1f | %.3f | %s | %s | %s | %s | %s | %s | %s | %s | %s |" % (
            r["kernel"], ("%g" % round(r["launches"], 2)), r["avg_us"], r["total_ms"],
            r.get("mfma_type") or "", f("mfma_gflop_per_launch", "%.3f"), f("mfma_tflops", "%.2f"),
            f("frac_peak", "%.3f"), f("mfma_util_pct", "%.1f"), f("read_mb", "%.2f"),
            f("write_mb", "%.2f"), f("hbm_gbs", "%.0f"), f("frac_hbm", "%.3f")))
    return "\n".join(lines)


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--mfma")
    ap.add_argument("--fetch")
    ap.add_argument("--write")
    ap.add_argument("--range", default="0:999999999", help="first:last dispatch id")
    ap.add_argument("--per", type=float, default=1.0, help="divide launches and totals by this")
    ap.add_argument("--copy-bytes", type=float, default=None,
                    help="bytes each way of a device copy in the run, for the FETCH/WRITE scale")
    ap.add_argument("--top", type=int, default=25)
    a = ap.parse_args(argv)
    mfma = load(a.mfma) if a.mfma else None
    fetch = load(a.fetch) if a.fetch else None
    write = load(a.write) if a.write else None
    fscale = wscale = 1024.0
    if fetch is not None and write is not None and a.copy_bytes:
        fscale, wscale = copy_scales(fetch, write, a.copy_bytes)
        print(f"FETCH_SIZE scale {fscale / 1024:.4f} x KiB, WRITE_SIZE scale {wscale / 1024:.4f} x KiB",
              file=sys.stderr)
    lo, hi = (int(x) for x in a.range.split(":"))
    print(markdown(summarise(mfma, fetch, write, lo, hi, a.per, fscale, wscale), a.top))
    return 0


if __name__ == "__main__":
    sys.exit(main())
