"""Host-side checks of plans from ASTRA-like views (cmi_plan2d_create_vec, FanGeometry.from_vec
/ to_vec) that need no GPU: argument validation before any device call and the Python round
trip. The projections themselves are checked against the oracle in test_gpu_vecgeom.py and the
oracle's vector form in test_oracle_2d.py."""
import ctypes as C

import numpy as np
import pytest
import torch

import curadon_amd as cm
from curadon_amd import _lib


def _create(vecs, D=11, shape=(8, 8)):
    v = np.ascontiguousarray(np.asarray(vecs, dtype=np.float64))
    vs = np.array(shape, dtype=np.uint64)
    sp = np.ones(2)
    of = np.zeros(2)
    h = C.c_void_p()
    st = _lib._lib.cmi_plan2d_create_vec(
        0, 32, vs.ctypes.data_as(_lib._u64p), sp.ctypes.data_as(_lib._dp),
        of.ctypes.data_as(_lib._dp), 32, D, v.ctypes.data_as(_lib._dp), len(v), C.byref(h))
    return st, h


GOOD = [[0.0, 30.0, 0.0, -20.0, 1.0, 0.0], [30.0, 0.0, -20.0, 0.0, 0.0, -1.2]]


@pytest.mark.parametrize("bad", [
    [[0.0, 30.0, 0.0, -20.0, 0.0, 0.0]],           # zero cell step
    [[0.0, 30.0, 0.0, -20.0, 0.0, 1.0]],           # source on the detector line
    [[0.0, float("nan"), 0.0, -20.0, 1.0, 0.0]],   # non-finite
])
def test_bad_views_rejected(bad):
    st, h = _create(bad)
    assert st == _lib.ERR_INVALID_ARGUMENT, _lib.last_error()
    assert h.value is None
    assert "plan2d_create_vec" in _lib.last_error()


def test_null_views_rejected():
    assert _lib._lib.cmi_plan2d_create_vec(0, 32, None, None, None, 32, 7, None, 2,
                                           None) == _lib.ERR_INVALID_ARGUMENT


@pytest.mark.skipif(torch.cuda.is_available(), reason="checks the no-GPU path")
def test_valid_views_without_gpu_report_device_error():
    st, h = _create(GOOD)
    assert st == _lib.ERR_DEVICE, _lib.last_error()


def test_from_vec_keeps_views_and_round_trips():
    g = cm.FanGeometry(DSD=200.0, DSO=130.0, angles=np.linspace(0, 6, 5), vol_shape=(8, 8),
                       det_shape=11, det_spacing=1.3, det_offset=0.5)
    v = cm.FanGeometry.from_vec((8, 8), 11, g.to_vec(), vol_spacing=(1.0, 2.0))
    np.testing.assert_allclose(v.vecs, np.array(g.to_vec()))
    np.testing.assert_allclose(v.to_vec(), g.to_vec())
    assert v.det_spacing == pytest.approx(1.3)
    # the summary angle is that of source - detector centre: the view angle without a shift
    g0 = cm.FanGeometry(DSD=200.0, DSO=130.0, angles=np.linspace(0, 6, 5), vol_shape=(8, 8),
                        det_shape=11)
    v0 = cm.FanGeometry.from_vec((8, 8), 11, g0.to_vec())
    np.testing.assert_allclose(v0.angles, g0.angles % (2 * np.pi), atol=1e-12)
    np.testing.assert_allclose(v.vol_spacing, [1.0, 2.0])
    # the source shift of det_offset and the cell spacing appear in the views
    np.testing.assert_allclose(np.linalg.norm(np.array(g.to_vec())[:, 4:6], axis=1), 1.3)
    np.testing.assert_allclose(np.linalg.norm(np.array(g.to_vec())[:, 0:2], axis=1),
                               np.hypot(130.0, 0.5))


def test_unequal_spacings_accepted():
    v = cm.FanGeometry.from_vec((8, 8), 11, GOOD)
    assert v.vecs.shape == (2, 6)
    with pytest.raises(AttributeError):
        cm.FanGeometry.from_vec((8, 8), 11, [[0.0, 30.0, 0.0, -20.0, 0.0, 0.0]])
