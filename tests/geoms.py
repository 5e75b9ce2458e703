"""Geometry recipes shared by the tests (the benchmark workloads, scaled versions and tiny
cases). Plain records: the oracle reads the same attributes as curadon_amd.FanGeometry."""
from types import SimpleNamespace

import numpy as np


def fan(n, D, A, dso, dsd, du=1.0, det_offset=0.0, spacing=(1.0, 1.0), offset=(0.0, 0.0),
        arc=2 * np.pi, angles=None, ny=None):
    ny = n if ny is None else ny
    ang = np.linspace(0, arc, A, endpoint=False) if angles is None else np.asarray(angles, float)
    return SimpleNamespace(
        vol_shape=np.array([ny, n]), vol_spacing=np.array(spacing, float),
        vol_offset=np.array(offset, float), det_shape=int(D), det_spacing=float(du),
        det_offset=float(det_offset), angles=ang, DSO=np.full(ang.size, float(dso)),
        DSD=np.full(ang.size, float(dsd)), nangles=ang.size)


def c1():
    """64^2, 91 cells, 360 angles, DSO 128, DSD 192."""
    return fan(64, 91, 360, 128.0, 192.0)


def c2():
    """1024^2, 1449 cells, 720 angles, DSO 2048, DSD 3072."""
    return fan(1024, 1449, 720, 2048.0, 3072.0)


def c4():
    """512^2, 724 cells, 512 angles; the workload leaves DSO/DSD open: DSO = 2n, DSD = 3n."""
    return fan(512, 724, 512, 1024.0, 1536.0)


def tiny():
    """8x8 image, 7 cells, 12 angles: the golden-fixture case."""
    return fan(8, 7, 12, 16.0, 24.0)


def as_fan_geometry(rec, **kw):
    import curadon_amd as cm
    return cm.FanGeometry(DSD=rec.DSD, DSO=rec.DSO, angles=rec.angles,
                          vol_shape=rec.vol_shape, det_shape=rec.det_shape,
                          det_spacing=rec.det_spacing, det_offset=rec.det_offset,
                          vol_spacing=rec.vol_spacing, vol_offset=rec.vol_offset, **kw)


def cone(n, U, V, A, dso, dsd, du=(1.0, 1.0), angles=None, vol_spacing=(1.0, 1.0, 1.0),
         vol_offset=(0.0, 0.0, 0.0), det_offset=(0.0, 0.0), det_rotation=(0.0, 0.0, 0.0), cor=0.0,
         nz=None, ny=None):
    """Cone-beam record in the Python axis order of ConeGeometry: vol_shape (z, y, x),
    det_shape (v, u), angles (3, A) = (phi, theta, psi)."""
    nz = n if nz is None else nz
    ny = n if ny is None else ny
    if angles is None:
        angles = np.linspace(0, 2 * np.pi, A, endpoint=False)
    ang = np.asarray(angles, float)
    if ang.ndim == 1:
        ang = np.vstack([ang, np.zeros_like(ang), np.zeros_like(ang)])
    return SimpleNamespace(
        vol_shape=np.array([nz, ny, n]), vol_spacing=np.array(vol_spacing, float),
        vol_offset=np.array(vol_offset, float), det_shape=np.array([V, U]),
        det_spacing=np.array(du, float), det_offset=np.array(det_offset, float),
        det_rotation=np.array(det_rotation, float), angles=ang, DSO=float(dso), DSD=float(dsd),
        COR=float(cor), nangles=ang.shape[1])


def c3():
    """256^3 volume, 384 x 384 panel, 360 angles, DSO 512, DSD 768."""
    return cone(256, 384, 384, 360, 512.0, 768.0)


def as_cone_geometry(rec, **kw):
    import curadon_amd as cm
    return cm.ConeGeometry(DSD=rec.DSD, DSO=rec.DSO, angles=rec.angles, vol_shape=rec.vol_shape,
                           det_shape=rec.det_shape, det_spacing=rec.det_spacing,
                           det_offset=rec.det_offset, det_rotation=rec.det_rotation,
                           vol_spacing=rec.vol_spacing, vol_offset=rec.vol_offset, COR=rec.COR,
                           **kw)


def with_views(g, vecs):
    """A copy of record `g` with ASTRA-like views (the oracle then uses them)."""
    r = SimpleNamespace(**vars(g))
    r.vecs = np.asarray(vecs, dtype=np.float64)
    return r


def helix_views(n_views, n, radius, pitch_per_turn, dod, U, V, du=1.0, dv=1.0, tilt=0.0):
    """Helical views: the source on a helix about the y axis, the panel facing it at distance
    radius + dod, tilted by `tilt` about its own normal; an explicit, non-circular trajectory."""
    out = []
    for k in range(n_views):
        t = 2 * np.pi * k / n_views * 1.3
        h = pitch_per_turn * t / (2 * np.pi) - 0.3 * n
        s = np.array([-radius * np.sin(t), h, -radius * np.cos(t)])
        d = np.array([np.sin(t), 0.0, np.cos(t)])      # from the source towards the axis
        c = s + (radius + dod) * d
        e_u = np.array([np.cos(t), 0.0, -np.sin(t)])
        e_v = np.array([0.0, 1.0, 0.0])
        ct, st = np.cos(tilt), np.sin(tilt)
        u = du * (ct * e_u + st * e_v)
        v = dv * (-st * e_u + ct * e_v)
        out.append(np.concatenate([s, c, u, v]))
    return np.array(out)
