"""Host-side checks of the geometry descriptors that need no GPU: scalar arguments given as
one-element arrays (no NumPy deprecation), plan fingerprints that follow every attribute a plan
is built from, host-side backprojection matrices that follow a changed geometry, and the FBP
filter's full-FP32 product leaving the caller's matmul precision settings as they were."""
import warnings

import numpy as np
import pytest
import torch

import curadon_amd as cm
from curadon_amd import reconstruction


def _fan(**kw):
    args = dict(DSD=300.0, DSO=200.0, angles=np.linspace(0, 2 * np.pi, 12, endpoint=False),
                vol_shape=(16, 16), det_shape=23)
    args.update(kw)
    return cm.FanGeometry(**args)


def _cone(**kw):
    args = dict(DSD=300.0, DSO=200.0, angles=np.linspace(0, 2 * np.pi, 8, endpoint=False),
                vol_shape=(8, 10, 12), det_shape=(9, 13))
    args.update(kw)
    return cm.ConeGeometry(**args)


def test_one_element_arrays_without_deprecation():
    with warnings.catch_warnings():
        warnings.simplefilter("error")
        g = _fan(DSD=np.array([300.0]), DSO=np.array([200.0]), det_shape=np.array([23]),
                 det_spacing=np.array([1.5]), det_offset=np.array([0.25]),
                 det_rotation=np.array([0.0]))
        c = _cone(DSD=np.array([300.0]), DSO=np.array([200.0]), COR=np.array([0.5]))
    assert np.all(g.DSD == 300.0) and np.all(g.DSO == 200.0) and g.det_shape == 23
    assert g.det_spacing == 1.5 and g.det_offset == 0.25
    assert c.DSD == 300.0 and c.DSO == 200.0 and c.COR == 0.5


@pytest.mark.parametrize("change", [
    lambda g: setattr(g, "angles", g.angles + 0.1),
    lambda g: g.angles.__setitem__(3, 1.0),          # in-place change of the array
    lambda g: setattr(g, "det_offset", 0.5),
    lambda g: setattr(g, "det_spacing", 2.0),
    lambda g: g.DSO.__setitem__(0, 210.0),
    lambda g: setattr(g, "vol_offset", np.array([1.0, 0.0])),
    lambda g: setattr(g, "sino_prec", 16),
])
def test_fan_fingerprint_follows_attributes(change):
    g = _fan()
    before = g.fingerprint()
    assert g.fingerprint() == before
    change(g)
    assert g.fingerprint() != before


@pytest.mark.parametrize("change", [
    lambda g: g.angles.__setitem__((0, 2), 0.3),
    lambda g: setattr(g, "COR", 1.0),
    lambda g: setattr(g, "DSO", 250.0),
    lambda g: g.det_offset.__setitem__(1, 0.5),
    lambda g: setattr(g, "det_rotation", np.array([0.0, 0.1, 0.0])),
])
def test_cone_fingerprint_follows_attributes(change):
    g = _cone()
    before = g.fingerprint()
    change(g)
    assert g.fingerprint() != before


def test_cone_host_matrices_follow_geometry():
    g = _cone()
    m0, yinv = g.bp_matrices("fdk", with_layout=True)
    assert m0.shape == (8, 3, 4) and m0.dtype == np.float32 and yinv
    g.angles[0, 1] += 0.25
    m1 = g.bp_matrices("fdk")
    assert not np.array_equal(m0[1], m1[1]) and np.array_equal(m0[2:], m1[2:])
    # the voxel at the volume centre projects onto the panel centre in matched mode:
    # U / W and V / W at h = (0, 0, 0, 1) are the centre pixel coordinates (nu - 1) / 2 ...
    m = _cone().bp_matrices("matched")
    u = m[:, 0, 3] / m[:, 2, 3]
    v = m[:, 1, 3] / m[:, 2, 3]
    assert np.allclose(u, (13 - 1) / 2, atol=1e-4) and np.allclose(v, (9 - 1) / 2, atol=1e-4)


@pytest.mark.parametrize("setting", ["default", "medium", "legacy", "ieee"])
def test_full_fp32_product_restores_precision_settings(setting):
    m = torch.backends.cuda.matmul
    saved = m.fp32_precision
    try:
        if setting == "medium":
            torch.set_float32_matmul_precision("medium")
        elif setting == "legacy":
            m.allow_tf32 = True
        elif setting == "ieee":
            m.fp32_precision = "ieee"

        def state():
            return (torch.get_float32_matmul_precision(), m.allow_tf32, m.fp32_precision)
        before = state()
        g = torch.Generator().manual_seed(1)
        a = torch.rand((5, 7), generator=g)
        b = torch.rand((7, 3), generator=g)
        y = reconstruction._mm_ieee(a, b)
        assert state() == before
        assert torch.allclose(y.double(), a.double() @ b.double(), rtol=1e-6, atol=0)
        if setting == "medium":
            assert torch.get_float32_matmul_precision() == "medium"
    finally:
        torch.set_float32_matmul_precision("highest")
        m.fp32_precision = saved
