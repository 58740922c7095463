"""The rocprofv3 counter summariser on small hand-made CSVs (CPU only)."""
import csv

import pytest

from mat_amd import rocprof_summary as rs

HEAD = ["Correlation_Id", "Dispatch_Id", "Agent_Id", "Queue_Id", "Process_Id", "Thread_Id",
        "Grid_Size", "Kernel_Id", "Kernel_Name", "Workgroup_Size", "LDS_Block_Size",
        "Scratch_Size", "VGPR_Count", "Accum_VGPR_Count", "SGPR_Count", "Counter_Name",
        "Counter_Value", "Start_Timestamp", "End_Timestamp"]

GEMM = "void mqr::(anonymous namespace)::k_gemm_f64<128, false, false, 0, 0>(mqr::GemmF64, int)"
COPY = "__amd_rocclr_copyBuffer"


def write(path, rows):
    with open(path, "w", newline="") as f:
        w = csv.writer(f)
        w.writerow(HEAD)
        for d, name, grid, ctr, val, t0, t1 in rows:
            w.writerow([d, d, "Agent 1", 1, 1, 1, grid, 1, name, 256, 0, 0, 8, 0, 8, ctr, val, t0, t1])


def test_summary_scales_and_splits_by_k(tmp_path):
    copy = 1 << 20  # 1 MiB copied each way
    # two update launches with 16 tiles of 128x128: K = 64 and K = 256
    flops64 = 2 * 16 * 128 * 128 * 64
    flops256 = 4 * flops64
    mfma = [(1, COPY, 1024, "SQ_INSTS_VALU_MFMA_MOPS_F64", 0, 0, 100),
            (2, GEMM, 16 * 256, "SQ_INSTS_VALU_MFMA_MOPS_F64", flops64 / 512, 0, 1000),
            (3, GEMM, 16 * 256, "SQ_INSTS_VALU_MFMA_MOPS_F64", flops256 / 512, 0, 2000)]
    fetch = [(1, COPY, 1024, "FETCH_SIZE", copy / 2048, 0, 100),   # counter reads half
             (2, GEMM, 16 * 256, "FETCH_SIZE", 100.0, 0, 1000),
             (3, GEMM, 16 * 256, "FETCH_SIZE", 200.0, 0, 2000)]
    write_ = [(1, COPY, 1024, "WRITE_SIZE", copy / 1024, 0, 100),
              (2, GEMM, 16 * 256, "WRITE_SIZE", 50.0, 0, 1000),
              (3, GEMM, 16 * 256, "WRITE_SIZE", 50.0, 0, 2000)]
    for name, rows in (("m", mfma), ("f", fetch), ("w", write_)):
        write(tmp_path / f"{name}.csv", rows)
    M, F, W = (rs.load(str(tmp_path / f"{n}.csv")) for n in "mfw")
    fs, ws = rs.copy_scales(F, W, copy)
    assert fs == pytest.approx(2048.0) and ws == pytest.approx(1024.0)
    recs = {r["kernel"]: r for r in rs.summarise(M, F, W, 2, 3, 1.0, fs, ws)}
    k64 = recs["k_gemm_f64<128, false, false, 0, 0> [K~64]"]
    k256 = recs["k_gemm_f64<128, false, false, 0, 0> [K~256]"]
    assert k64["launches"] == 1 and k256["launches"] == 1
    assert k64["mfma_type"] == "F64"
    assert k64["mfma_tflops"] == pytest.approx(flops64 / 1e-6 / 1e12)
    assert k256["read_mb"] == pytest.approx(200 * 2048 / 1e6)
    assert k256["write_mb"] == pytest.approx(50 * 1024 / 1e6)
    assert k256["hbm_gbs"] == pytest.approx((200 * 2048 + 50 * 1024) / 2e-6 / 1e9)


def test_summary_rejects_mismatched_passes(tmp_path):
    write(tmp_path / "m.csv", [(1, GEMM, 256, "SQ_INSTS_VALU_MFMA_MOPS_F64", 1, 0, 10)])
    write(tmp_path / "f.csv", [(1, COPY, 256, "FETCH_SIZE", 1, 0, 10)])
    M, F = rs.load(str(tmp_path / "m.csv")), rs.load(str(tmp_path / "f.csv"))
    with pytest.raises(SystemExit):
        rs.summarise(M, F, None, 0, 10, 1.0, 1024.0, 1024.0)
