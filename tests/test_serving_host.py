"""CPU checks of the serving ABI's host helpers: the per-angle tables a serving job carries
(against closed forms and the FP64 oracle), tile counts, and argument validation."""
import ctypes as C
import math

import numpy as np
import pytest

from curadon_amd import _lib
from geoms import cone


def test_fan2d_bp_table_closed_form():
    ang = np.linspace(0, 2 * np.pi, 37, endpoint=False) + 0.1
    dso = np.linspace(100, 140, 37)
    dsd = dso * 1.5
    t = _lib.fan2d_bp_table(ang, dso, dsd)
    ref = np.stack([-np.sin(ang), np.cos(ang), dsd, dso], axis=1)
    np.testing.assert_allclose(t, ref, rtol=2e-7, atol=1e-7)
    assert t.dtype == np.float32


def _plan3d(rec, prec=32):
    """A plan_3d object without a device: only its descriptor is used here."""
    p = _lib.plan_3d.__new__(_lib.plan_3d)
    p._h = None
    ang = rec.angles
    p._phi, p._theta, p._psi = (np.ascontiguousarray(ang[i], dtype=np.float64) for i in range(3))
    d = _lib.Cone3DDesc()
    d.vol_shape[:] = [int(v) for v in rec.vol_shape]
    d.vol_spacing[:] = list(rec.vol_spacing)
    d.vol_offset[:] = list(rec.vol_offset)
    d.det_shape[:] = [int(v) for v in rec.det_shape]
    d.det_spacing[:] = list(rec.det_spacing)
    d.det_offset[:] = list(rec.det_offset)
    d.det_rotation[:] = list(rec.det_rotation)
    d.DSO, d.DSD, d.COR = rec.DSO, rec.DSD, rec.COR
    d.nangles = ang.shape[1]
    d.phi = p._phi.ctypes.data_as(_lib._dp)
    d.theta = p._theta.ctypes.data_as(_lib._dp)
    d.psi = p._psi.ctypes.data_as(_lib._dp)
    p._desc = d
    p.nangles = ang.shape[1]
    return p


def _bp_with_matrices(M, sino, vol_shape, mode):
    """Backprojection the serving kernel computes from the 3x4 matrices, in FP64: lower-tap
    panel coordinate (U/W, V/W), bilinear with zeros off the panel; 'matched' / 'fdk' drop
    W <= 0, 'fdk' weights by 1 / W^2."""
    nz, ny, nx = vol_shape
    A, V, U = sino.shape
    c = [(n - 1) / 2.0 for n in (nx, ny, nz)]
    z, y, x = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    h = np.stack([x - c[0], y - c[1], z - c[2], np.ones_like(x, dtype=float)], axis=-1)
    pad = np.zeros((A, V + 4, U + 4))
    pad[:, 2:-2, 2:-2] = sino
    out = np.zeros(vol_shape)
    for a in range(A):
        Uh, Vh, Wh = np.moveaxis(h @ M[a].astype(np.float64).T, -1, 0)
        u = np.clip(Uh / Wh, -2.0, U)
        v = np.clip(Vh / Wh, -2.0, V)
        iu, iv = np.floor(u).astype(int), np.floor(v).astype(int)
        fu, fv = u - iu, v - iv
        iu, iv = iu + 2, iv + 2
        P = pad[a]
        s = ((1 - fv) * ((1 - fu) * P[iv, iu] + fu * P[iv, iu + 1]) +
             fv * ((1 - fu) * P[iv + 1, iu] + fu * P[iv + 1, iu + 1]))
        if mode == "fdk":
            s = s / Wh ** 2
        if mode in ("matched", "fdk"):
            s = np.where(Wh > 0, s, 0.0)
        out += s
    return out


@pytest.mark.parametrize("mode", ["reference", "matched", "fdk"])
def test_cone3d_bp_matrices_against_oracle(oracle, mode):
    r = np.random.default_rng(1)
    ang = np.vstack([r.uniform(0, 2 * np.pi, 6), r.uniform(-0.2, 0.2, 6),
                     r.uniform(-0.2, 0.2, 6)])
    rec = cone(10, 14, 12, 6, 30.0, 50.0, du=(1.1, 0.9), angles=ang,
               vol_spacing=(1.0, 1.2, 0.9), vol_offset=(0.3, -0.4, 0.2), det_offset=(0.5, -0.7),
               det_rotation=(0.02, -0.03, 0.01), cor=0.4, nz=8, ny=9)
    M = _plan3d(rec).bp_matrices(mode)
    assert M.shape == (6, 3, 4) and M.dtype == np.float32
    y = r.random((6, 12, 14))
    got = _bp_with_matrices(M, y, (8, 9, 10), mode)
    ref = oracle.bp3d(rec, y, matched=mode == "matched", fdk=mode == "fdk")
    err = np.linalg.norm(got - ref) / np.linalg.norm(ref)
    print(f"bp matrices ({mode}) vs FP64 oracle: rel-L2 {err:.3e}")
    assert err <= 2e-5


def test_job_tiles():
    j = _lib.BPJob(kind=_lib.JOB_FAN2D, nx=130, ny=9, nz=1)
    assert _lib.bp_job_tiles(j) == 3 * 3
    j = _lib.BPJob(kind=_lib.JOB_CONE3D, nx=33, ny=16, nz=17)
    assert _lib.bp_job_tiles(j) == 3 * 1 * 3
    with pytest.raises(ValueError):
        _lib.bp_job_tiles(_lib.BPJob(kind=7, nx=1, ny=1, nz=1))


def test_stage_and_server_argument_errors():
    st = _lib._lib.cmi_stage_projections(0, None, 8, 2, 1, 8, 0, None, 1.0, None, _lib.F16, None)
    assert st == _lib.ERR_INVALID_ARGUMENT and "null" in _lib.last_error()
    st = _lib._lib.cmi_stage_projections(0, 16, 8, 2, 1, 8, 0, None, 1.0, 32, 77, None)
    assert st == _lib.ERR_DTYPE
    st = _lib._lib.cmi_stage_projections(0, 16, 4, 2, 1, 8, 0, None, 1.0, 32, _lib.F16, None)
    assert st == _lib.ERR_INVALID_ARGUMENT  # stride shorter than a row
    h = C.c_void_p()
    assert _lib._lib.cmi_server_create(0, 0, C.byref(h)) == _lib.ERR_INVALID_ARGUMENT
    assert _lib._lib.cmi_serve_backproject(None, None, 0, 32, 32, None) == \
        _lib.ERR_INVALID_ARGUMENT
    assert _lib._lib.cmi_server_destroy(None) == _lib.OK


def test_fan2d_bp_table_matches_plan_formula_for_python_floats():
    # the serving descriptor passes 1 / du rounded to FP32, as the plan's launcher does
    du = 0.7
    assert np.float32(1.0 / du) == C.c_float(1.0 / du).value
    assert math.isclose(_lib.fan2d_bp_table([0.5], [10.0], [15.0])[0, 2], 15.0)


def test_y_invariance_flag():
    """Circular orbits about the y axis leave U and W free of y in every mode, also with the
    panel turned about y (pitch); a rolled panel or an out-of-plane angle does not."""
    circ = cone(10, 14, 12, 6, 30.0, 50.0, det_offset=(0.5, -0.7), cor=0.4)
    for mode in ("reference", "matched", "fdk"):
        M, yinv = _plan3d(circ).bp_matrices(mode, with_layout=True)
        assert yinv, mode
        assert np.all(M[:, 0, 1] == 0) and np.all(M[:, 2, 1] == 0)
    pitched = cone(10, 14, 12, 6, 30.0, 50.0, det_rotation=(0.02, 0.0, 0.0))
    assert _plan3d(pitched).bp_matrices("fdk", with_layout=True)[1]
    rolled = cone(10, 14, 12, 6, 30.0, 50.0, det_rotation=(0.0, 0.02, 0.0))
    assert not _plan3d(rolled).bp_matrices("fdk", with_layout=True)[1]
    ang = np.vstack([np.linspace(0, 6, 6), np.full(6, 0.1), np.zeros(6)])
    oblique = cone(10, 14, 12, 6, 30.0, 50.0, angles=ang)
    assert not _plan3d(oblique).bp_matrices("matched", with_layout=True)[1]


def test_job_tiles_y_runs():
    # tiles of 16 (x) x 16 (y) x 16 (z) voxels
    j = _lib.BPJob(kind=_lib.JOB_CONE3D, columns=1, nx=33, ny=16, nz=17)
    assert _lib.bp_job_tiles(j) == 3 * 1 * 2
    j = _lib.BPJob(kind=_lib.JOB_CONE3D, columns=1, nx=16, ny=17, nz=48)
    assert _lib.bp_job_tiles(j) == 1 * 2 * 3
