"""CPU checks of the C-ABI library: it loads, exports every symbol the public header
declares, and its host-side validation rejects bad arguments before touching a GPU."""
import ctypes as C
import os
import re

import numpy as np
import pytest

from curadon_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    names = set()
    for f in os.listdir(os.path.join(ROOT, "include")):
        if f.endswith(".h"):
            txt = open(os.path.join(ROOT, "include", f)).read()
            names |= set(re.findall(r"CMI_API\s+[\w\s\*]+?\b(cmi_\w+)\s*\(", txt))
    return names


def test_every_declared_symbol_is_exported():
    names = declared_symbols()
    assert len(names) >= 15
    lib = C.CDLL(_lib.LIB_PATH)
    for n in sorted(names):
        assert hasattr(lib, n), n
    assert names == set(_lib.SIGNATURES), names ^ set(_lib.SIGNATURES)


def test_version_and_error_string():
    assert _lib.version() >= 100
    assert isinstance(_lib.last_error(), str)


def _plan(**over):
    args = dict(device=0, vol_prec=32, shape=(8, 8), spacing=(1.0, 1.0), offset=(0.0, 0.0),
                det_prec=32, det=7, du=1.0, doff=0.0, A=4)
    args.update(over)
    A = args["A"]
    ang = np.linspace(0, 1, max(A, 1))
    dso = np.full(max(A, 1), 16.0)
    dsd = np.full(max(A, 1), 24.0)
    vs = np.array(args["shape"], dtype=np.uint64)
    sp = np.array(args["spacing"], dtype=np.float64)
    of = np.array(args["offset"], dtype=np.float64)
    h = C.c_void_p()
    dp = C.POINTER(C.c_double)
    st = _lib._lib.cmi_plan2d_create(
        args["device"], args["vol_prec"], vs.ctypes.data_as(C.POINTER(C.c_uint64)),
        sp.ctypes.data_as(dp), of.ctypes.data_as(dp), args["det_prec"], args["det"],
        args["du"], args["doff"], dso.ctypes.data_as(dp), dsd.ctypes.data_as(dp),
        ang.ctypes.data_as(dp), A, 0.0, 0.0, C.byref(h))
    return st, h


@pytest.mark.parametrize("over,status", [
    (dict(vol_prec=8), _lib.ERR_DTYPE),
    (dict(det_prec=64), _lib.ERR_DTYPE),
    (dict(shape=(0, 8)), _lib.ERR_INVALID_ARGUMENT),
    (dict(det=0), _lib.ERR_INVALID_ARGUMENT),
    (dict(A=0), _lib.ERR_INVALID_ARGUMENT),
    (dict(spacing=(1.0, -1.0)), _lib.ERR_INVALID_ARGUMENT),
    (dict(du=0.0), _lib.ERR_INVALID_ARGUMENT),
])
def test_plan_validation(over, status):
    st, h = _plan(**over)
    assert st == status, _lib.last_error()
    assert h.value is None
    assert _lib.last_error()


def test_null_arguments():
    assert _lib._lib.cmi_plan2d_create(0, 32, None, None, None, 32, 7, 1.0, 0.0, None, None,
                                       None, 4, 0.0, 0.0, None) == _lib.ERR_INVALID_ARGUMENT
    assert _lib._lib.cmi_forward_2d(None, None, 32, None, 32, 1, None) == _lib.ERR_INVALID_ARGUMENT
    assert "null" in _lib.last_error()
    assert _lib._lib.cmi_backward_2d(None, None, 32, None, 32, 1, None) == _lib.ERR_INVALID_ARGUMENT
    assert _lib._lib.cmi_adjoint_2d(None, None, 32, None, 32, 1, None) == _lib.ERR_INVALID_ARGUMENT
    assert _lib._lib.cmi_plan2d_destroy(None) == _lib.OK
    assert _lib._lib.cmi_plan2d_set_block(None, 64, 4) == _lib.ERR_INVALID_ARGUMENT


@pytest.mark.skipif(__import__("torch").cuda.is_available(), reason="checks the no-GPU path")
def test_valid_plan_without_gpu_reports_device_error():
    st, h = _plan()
    assert st == _lib.ERR_DEVICE, _lib.last_error()
    assert "device" in _lib.last_error()


def test_precision_codes_match_reference():
    assert _lib.precision_code(32) == _lib.F32
    assert _lib.precision_code(16) == _lib.F16
    assert _lib.precision_code("bf16") == _lib.BF16
    with pytest.raises(AttributeError, match="Precision must"):
        _lib.precision_code(8)
