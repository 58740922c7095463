"""The header-only C++ drop-in (include/mat/Mat.h): the reference's QR() /
QR_init() call shapes and LeastSquares, compiled with g++ against
libmatqr.so (tests/cpp/test_mat_api.cpp), run on the GPU."""
import os
import subprocess

import pytest

import mat_amd as M

pytestmark = pytest.mark.gpu
BIN = os.path.join(os.path.dirname(M.LIB_PATH), "test_mat_api")


def test_cpp_header_api():
    assert os.path.exists(BIN), "build the library first (python -m mat_amd.build)"
    r = subprocess.run([BIN], capture_output=True, text=True, timeout=100)
    print(r.stdout)
    print(r.stderr)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "ALL PASSED" in r.stdout
