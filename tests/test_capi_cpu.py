"""CPU-side checks of the C ABI library: it loads, exports every entry point
include/mat_qr.h declares, and reports the absence of a device with a defined
status instead of crashing.  No compute call is made here (there is no GPU in
this environment)."""
import ctypes as C
import os
import re

import pytest

import mat_amd

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared_symbols():
    text = open(os.path.join(ROOT, "include", "mat_qr.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(mqr_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    L = mat_amd.lib()
    declared = _declared_symbols()
    assert len(declared) >= 20
    for name in declared:
        assert hasattr(L, name), name
    assert set(declared) == set(mat_amd.SIGNATURES), "Python bindings out of sync with the header"


def test_strerror_is_defined_for_every_code():
    for code in [0, -1, -7, 1001, 1002, 1003, 1004, 1005, 1006, 1007, 4242]:
        s = mat_amd.strerror(code)
        assert isinstance(s, str) and s


def test_null_handle_is_argument_one():
    L = mat_amd.lib()
    assert L.mqr_dgeqrf(None, 4, 4, None, 4, None, None, 64) == -1
    assert L.mqr_dgemm(None, b"N", b"N", 1, 1, 1, 1.0, None, 1, None, 1, 0.0, None, 1) == -1
    assert L.mqr_gels_mixed(None, 0, 4, 2, None, 4, None, None, None, None) == -1
    assert L.mqr_destroy(None) == -1
    assert L.mqr_create(None, 0) == -1


def test_no_device_is_a_defined_error():
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    L = mat_amd.lib()
    n = C.c_int(-1)
    assert L.mqr_get_device_count(C.byref(n)) == mat_amd.MQR_ERR_NO_DEVICE
    assert n.value == 0
    h = C.c_void_p()
    assert L.mqr_create(C.byref(h), 0) == mat_amd.MQR_ERR_NO_DEVICE
    assert not h.value


def test_product_does_not_reference_the_oracle():
    """The shipped package must not import, load or name the oracle."""
    pkg = os.path.join(ROOT, "mat_amd")
    for dirpath, _, files in os.walk(pkg):
        for f in files:
            if f.endswith((".py", ".hip", ".h", ".cpp")):
                text = open(os.path.join(dirpath, f)).read()
                assert "oracle" not in text.lower(), f
