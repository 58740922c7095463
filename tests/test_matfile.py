"""The binary MAT file format (Mat-5.2/_extra/__File.h:1-66) through the
library's readers and writers (host code; no GPU needed).  Layout pinned by
files written with numpy alone (tests/golden/make_golden.py): int32 rows,
int32 cols, then the elements column-major with no type tag."""
import os

import numpy as np
import pytest

import mat_amd as M

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLD = os.path.join(ROOT, "tests", "golden")


def test_reads_the_numpy_written_fixtures(golden):
    A = M.mat_load(os.path.join(GOLD, "A_17x9.mat"))
    VR = M.mat_load(os.path.join(GOLD, "VR_17x9.mat"))
    A32 = M.mat_load(os.path.join(GOLD, "A_17x9_f32.mat"), dtype="float32")
    assert A.shape == (17, 9) and A.flags.f_contiguous
    assert np.array_equal(A, golden["A_17x9"])
    assert np.array_equal(VR, golden["VR_17x9"])
    assert np.array_equal(A32, golden["A_17x9"].astype(np.float32))


def test_writes_byte_for_byte_what_the_format_specifies(tmp_path):
    rng = np.random.default_rng(3)
    for dt, shape in [("float64", (5, 3)), ("float32", (1, 7)), ("float64", (0, 4)),
                      ("int32", (3, 3))]:
        A = np.asfortranarray((rng.standard_normal(shape) * 100).astype(dt))
        p = str(tmp_path / f"x_{dt}_{shape[0]}x{shape[1]}.mat")
        M.mat_save(p, A)
        raw = open(p, "rb").read()
        assert raw[:8] == np.array(shape, dtype="<i4").tobytes()
        assert raw[8:] == A.tobytes(order="F")
        size = __import__("ctypes").c_int64(0)
        M.check(M.lib().mqr_file_size(p.encode(), __import__("ctypes").byref(size)), "size")
        assert size.value == 8 + A.size * A.itemsize
        assert np.array_equal(M.mat_load(p, dtype=dt), A)


def test_raw_forms_round_trip(tmp_path):
    import ctypes as C
    x = np.arange(11, dtype=np.float64) * 0.25
    p = str(tmp_path / "raw.bin").encode()
    L = M.lib()
    M.check(L.mqr_raw_save(p, x.ctypes.data_as(C.c_void_p), x.nbytes), "raw_save")
    y = np.zeros(11)
    M.check(L.mqr_raw_load(p, y.ctypes.data_as(C.c_void_p), y.nbytes), "raw_load")
    assert np.array_equal(x, y)
    # asking for more bytes than the file holds is an error, not a partial read
    z = np.zeros(12)
    assert L.mqr_raw_load(p, z.ctypes.data_as(C.c_void_p), z.nbytes) == M.MQR_ERR_IO


def test_errors_are_reported(tmp_path):
    import ctypes as C
    L = M.lib()
    missing = str(tmp_path / "missing.mat")
    with pytest.raises(M.MqrError) as e:
        M.mat_load(missing)
    assert e.value.status == M.MQR_ERR_IO
    p = str(tmp_path / "a.mat")
    A = np.asfortranarray(np.ones((4, 3)))
    M.mat_save(p, A)
    buf = np.zeros((4, 3), order="F")
    # header disagrees with the requested shape
    assert L.mqr_mat_load(p.encode(), 3, 4, buf.ctypes.data_as(C.c_void_p), 8) == M.MQR_ERR_IO
    # element size disagrees with the file length
    assert L.mqr_mat_load(p.encode(), 4, 3, buf.ctypes.data_as(C.c_void_p), 4) == M.MQR_ERR_IO
    # truncated file
    open(p, "r+b").truncate(8 + 8 * 11)
    assert L.mqr_mat_load(p.encode(), 4, 3, buf.ctypes.data_as(C.c_void_p), 8) == M.MQR_ERR_IO
    # LAPACK-style argument positions
    assert L.mqr_mat_save(None, 1, 1, None, 8) == -1
    assert L.mqr_mat_save(p.encode(), -1, 1, None, 8) == -2
    assert L.mqr_mat_save(p.encode(), 1, 1, None, 8) == -4
    assert L.mqr_mat_save(p.encode(), 1, 1, buf.ctypes.data_as(C.c_void_p), 3) == -5
