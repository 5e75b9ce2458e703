"""ctypes front end of the FP64 oracle (oracle/curadon_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, by ``__graft_entry__.smoke()`` and by the
CPU-baseline leg of ``bench.py``. The product package never imports it.

Geometry arguments mirror the product's geometry records (``FanGeometry``/``ConeGeometry``
attributes), so a test can hand the same record to both sides.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "libcuradon_oracle.so")
_lib = None

_i64 = C.c_int64
_dp = C.POINTER(C.c_double)
_fp = C.POINTER(C.c_float)


class Fan2D(C.Structure):
    _fields_ = [
        ("nx", _i64), ("ny", _i64), ("sx", C.c_double), ("sy", C.c_double),
        ("ox", C.c_double), ("oy", C.c_double), ("det_count", _i64), ("du", C.c_double),
        ("det_offset", C.c_double), ("nangles", _i64), ("angles", _dp), ("dso", _dp),
        ("dsd", _dp), ("vecs", _dp),
    ]


class Cone3D(C.Structure):
    _fields_ = [
        ("nx", _i64), ("ny", _i64), ("nz", _i64), ("sx", C.c_double), ("sy", C.c_double),
        ("sz", C.c_double), ("ox", C.c_double), ("oy", C.c_double), ("oz", C.c_double),
        ("nu", _i64), ("nv", _i64), ("du", C.c_double), ("dv", C.c_double),
        ("off_u", C.c_double), ("off_v", C.c_double), ("pitch", C.c_double),
        ("roll", C.c_double), ("yaw", C.c_double), ("dso", C.c_double), ("dsd", C.c_double),
        ("cor", C.c_double), ("nangles", _i64), ("phi", _dp), ("theta", _dp), ("psi", _dp),
        ("vecs", _dp),
    ]


def build() -> str:
    """Compile the oracle with its Makefile (gcc, no GPU needed)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = C.CDLL(_LIB_PATH)
        F2, F3 = C.POINTER(Fan2D), C.POINTER(Cone3D)
        sig = {
            "orc_plan2d": (None, [F2, _dp, _dp, _dp]),
            "orc_ray2d": (_i64, [F2, _dp, _dp, _dp, _i64, _i64, _dp, _dp, _dp]),
            "orc_fwd2d": (None, [F2, _dp, _dp, _i64, _i64, C.POINTER(_i64)]),
            "orc_adj2d": (None, [F2, _dp, _dp, _i64, _i64]),
            "orc_bp2d": (None, [F2, _dp, _dp, _i64, _i64]),
            "orc_bp2d_w": (None, [F2, _dp, _dp, _i64, _i64, C.c_int]),
            "orc_bp2d_T_w": (None, [F2, _dp, _dp, _i64, _i64, C.c_int]),
            "orc_bp3d_fdk": (None, [F3, _dp, _dp, _i64, _i64]),
            "orc_bp3d_T": (None, [F3, C.c_int, _dp, _dp, _i64, _i64]),
            "orc_fwd3d_f32": (None, [F3, _fp, _fp, _i64, _i64]),
            "orc_bp3d_f32": (None, [F3, _fp, _fp, _i64, _i64]),
            "orc_fwd2d_f32": (None, [F2, _fp, _fp, _i64, _i64]),
            "orc_bp2d_f32": (None, [F2, _fp, _fp, _i64, _i64]),
            "orc_num_threads": (C.c_int, []),
            "orc_set_num_threads": (None, [C.c_int]),
            "orc_lerp1": (C.c_double, [_dp, _i64, C.c_double]),
            "orc_bilerp": (C.c_double, [_dp, _i64, _i64, C.c_double, C.c_double]),
            "orc_trilerp": (C.c_double, [_dp, _i64, _i64, _i64] + [C.c_double] * 3),
            "orc_rotate_yzy": (None, [_dp] + [C.c_double] * 3 + [_dp]),
            "orc_rotate_rpy": (None, [_dp] + [C.c_double] * 3 + [_dp]),
            "orc_plan3d_fwd": (None, [F3, _dp, _dp, _dp, _dp]),
            "orc_ray3d": (_i64, [F3, _dp, _dp, _dp, _dp, _i64, _i64, _i64, _dp, _dp, _dp, _dp]),
            "orc_fwd3d": (None, [F3, _dp, _dp, _i64, _i64, C.POINTER(_i64)]),
            "orc_adj3d": (None, [F3, _dp, _dp, _i64, _i64]),
            "orc_plan3d_bp": (None, [F3, _dp, _dp, _dp, _dp, _dp]),
            "orc_bp3d": (None, [F3, _dp, _dp, _i64, _i64]),
            "orc_bp3d_matched": (None, [F3, _dp, _dp, _i64, _i64]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def num_threads() -> int:
    return int(lib().orc_num_threads())


def set_num_threads(n: int) -> None:
    lib().orc_set_num_threads(int(n))


def _d(a):
    return a.ctypes.data_as(_dp)


def _f(a):
    return a.ctypes.data_as(_fp)


def _c64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


# ------------------------------------------------------------------------------ 2D -----

def exact_views2d(geom):
    """Views [A][6] of a circular fan geometry with exact detector geometry: before the clockwise
    rotation by the angle (the reference's rotate, rotation.hpp:51-61), the source at
    (-COR, DSO), the detector centre at (det_offset - COR, -(DSD - DSO)) and the cell step
    det_spacing (cos r, sin r), r = det_rotation. The reference itself moves the source by
    det_offset in its forward only and leaves det_pitch / COR unused (plan_2d.hpp:76-80,
    121-122); this is the consistent geometry that the exact-geometry plans implement."""
    ang = np.asarray(geom.angles, dtype=np.float64)
    A = ang.size
    dso = np.broadcast_to(np.asarray(geom.DSO, dtype=np.float64), (A,))
    dsd = np.broadcast_to(np.asarray(geom.DSD, dtype=np.float64), (A,))
    cor = float(getattr(geom, "COR", 0.0))
    r = float(getattr(geom, "det_rotation", 0.0))
    du, off = float(geom.det_spacing), float(geom.det_offset)
    out = np.zeros((A, 6))
    for a in range(A):
        c, s = np.cos(ang[a]), np.sin(ang[a])

        def rot(x, y):
            return (x * c + y * s, -x * s + y * c)
        out[a, 0:2] = rot(-cor, dso[a])
        out[a, 2:4] = rot(off - cor, -(dsd[a] - dso[a]))
        out[a, 4:6] = rot(du * np.cos(r), du * np.sin(r))
    return out


class Fan:
    """Holds an ``orc_fan2d`` and the arrays it points into.

    ``geom`` is any record with the attributes of ``FanGeometry``: ``vol_shape`` (y, x),
    ``vol_spacing``/``vol_offset`` (y, x), ``det_shape``, ``det_spacing``, ``det_offset``,
    ``angles``, ``DSO``, ``DSD`` (per angle); with a ``vecs`` attribute ([A][6] ASTRA-like
    views) that is not None, the views define the geometry instead, and with a true
    ``exact_geometry`` attribute the views of ``exact_views2d``.
    """

    def __init__(self, geom):
        vs = np.asarray(geom.vol_shape)
        sp = np.asarray(geom.vol_spacing, dtype=np.float64)
        of = np.asarray(geom.vol_offset, dtype=np.float64)
        vecs = getattr(geom, "vecs", None)
        if vecs is None and getattr(geom, "exact_geometry", False):
            vecs = exact_views2d(geom)
        self.vecs = None if vecs is None else _c64(np.asarray(vecs, dtype=np.float64).reshape(-1, 6))
        self.angles = _c64(geom.angles) if self.vecs is None else _c64(np.zeros(len(self.vecs)))
        A = self.angles.size
        self.dso = _c64(np.broadcast_to(np.asarray(geom.DSO, dtype=np.float64), (A,)))
        self.dsd = _c64(np.broadcast_to(np.asarray(geom.DSD, dtype=np.float64), (A,)))
        self.s = Fan2D(
            nx=int(vs[1]), ny=int(vs[0]), sx=float(sp[1]), sy=float(sp[0]),
            ox=float(of[1]), oy=float(of[0]), det_count=int(geom.det_shape),
            du=float(geom.det_spacing), det_offset=float(geom.det_offset), nangles=A,
            angles=_d(self.angles), dso=_d(self.dso), dsd=_d(self.dsd),
            vecs=None if self.vecs is None else _d(self.vecs),
        )
        self.nx, self.ny, self.D, self.A = int(vs[1]), int(vs[0]), int(geom.det_shape), A

    def plan(self):
        src = np.zeros((self.A, 2))
        org = np.zeros((self.A, 2))
        dl = np.zeros((self.A, 2))
        lib().orc_plan2d(C.byref(self.s), _d(src), _d(org), _d(dl))
        return src, org, dl

    def ray(self, a, k, plan=None):
        src, org, dl = plan if plan is not None else self.plan()
        e = np.zeros(2)
        v = np.zeros(2)
        sc = np.zeros(1)
        n = lib().orc_ray2d(C.byref(self.s), _d(src), _d(org), _d(dl), a, k, _d(e), _d(v), _d(sc))
        return int(n), e, v, float(sc[0])


def fwd2d(geom, img, angles=None, count=False):
    """FP64 forward projection. ``angles`` = (a0, a1) restricts to those sinogram rows;
    the other rows are returned as NaN."""
    f = Fan(geom)
    img = _c64(img)
    assert img.shape == (f.ny, f.nx)
    a0, a1 = (0, f.A) if angles is None else angles
    sino = np.full((f.A, f.D), np.nan)
    ns = _i64(0)
    lib().orc_fwd2d(C.byref(f.s), _d(img), _d(sino), a0, a1, C.byref(ns))
    return (sino, int(ns.value)) if count else sino


def adj2d(geom, sino, angles=None):
    f = Fan(geom)
    sino = _c64(sino)
    assert sino.shape == (f.A, f.D)
    a0, a1 = (0, f.A) if angles is None else angles
    img = np.zeros((f.ny, f.nx))
    lib().orc_adj2d(C.byref(f.s), _d(sino), _d(img), a0, a1)
    return img


def bp2d(geom, sino, rows=None, img=None, wpow=1):
    """FP64 voxel-driven backprojection; adds into ``img`` (zeros if None). wpow = 2 uses the
    fan-beam FBP weight (squared magnification)."""
    f = Fan(geom)
    sino = _c64(sino)
    assert sino.shape == (f.A, f.D)
    out = np.zeros((f.ny, f.nx)) if img is None else _c64(img).copy()
    y0, y1 = (0, f.ny) if rows is None else rows
    lib().orc_bp2d_w(C.byref(f.s), _d(sino), _d(out), y0, y1, int(wpow))
    return out


def bp2d_T(geom, img, angles=None, wpow=1):
    """FP64 transpose of ``bp2d`` (the backprojection's weights applied to an image, into a
    sinogram); ``angles`` = (a0, a1) restricts it to those projections (the others stay 0)."""
    f = Fan(geom)
    img = _c64(img)
    assert img.shape == (f.ny, f.nx)
    out = np.zeros((f.A, f.D))
    a0, a1 = (0, f.A) if angles is None else angles
    lib().orc_bp2d_T_w(C.byref(f.s), _d(img), _d(out), a0, a1, int(wpow))
    return out


def fwd2d_f32(geom, img, angles=None):
    f = Fan(geom)
    img = np.ascontiguousarray(img, dtype=np.float32)
    a0, a1 = (0, f.A) if angles is None else angles
    sino = np.zeros((f.A, f.D), dtype=np.float32)
    lib().orc_fwd2d_f32(C.byref(f.s), _f(img), _f(sino), a0, a1)
    return sino


def bp2d_f32(geom, sino, rows=None):
    f = Fan(geom)
    sino = np.ascontiguousarray(sino, dtype=np.float32)
    out = np.zeros((f.ny, f.nx), dtype=np.float32)
    y0, y1 = (0, f.ny) if rows is None else rows
    lib().orc_bp2d_f32(C.byref(f.s), _f(sino), _f(out), y0, y1)
    return out


def lerp1(row, u):
    row = _c64(row)
    return lib().orc_lerp1(_d(row), row.size, float(u))


def bilerp(img, x, y):
    img = _c64(img)
    return lib().orc_bilerp(_d(img), img.shape[1], img.shape[0], float(x), float(y))


def trilerp(vol, x, y, z):
    vol = _c64(vol)
    nz, ny, nx = vol.shape
    return lib().orc_trilerp(_d(vol), nx, ny, nz, float(x), float(y), float(z))


# ------------------------------------------------------------------------------ 3D -----

def rotate_yzy(v, phi, theta, psi):
    v = _c64(v)
    out = np.zeros(3)
    lib().orc_rotate_yzy(_d(v), phi, theta, psi, _d(out))
    return out


def rotate_rpy(v, roll, pitch, yaw):
    v = _c64(v)
    out = np.zeros(3)
    lib().orc_rotate_rpy(_d(v), roll, pitch, yaw, _d(out))
    return out


class Cone:
    """Holds an ``orc_cone3d``. ``geom`` has ``ConeGeometry``'s attributes in Python axis
    order: ``vol_shape`` (z, y, x), ``det_shape`` (v, u), ``angles`` (3, A) as (phi, theta,
    psi), ``det_rotation`` (pitch, roll, yaw), scalar ``DSO``/``DSD``/``COR``. The axis
    reversal and the angle/rotation slots follow python/standalone/src/forward.cu:47-72.
    A ``vecs`` attribute (A, 12) that is not None (ASTRA-like views in world (x, y, z):
    source, panel centre, u-step, v-step) replaces the Euler-angle set-up; the other
    geometry fields are then unused apart from the shapes, spacings and offsets."""

    def __init__(self, geom):
        vs = np.asarray(geom.vol_shape)
        sp = np.asarray(geom.vol_spacing, dtype=np.float64)
        of = np.asarray(geom.vol_offset, dtype=np.float64)
        ds = np.asarray(geom.det_shape)
        vecs = getattr(geom, "vecs", None)
        self.vecs = None if vecs is None else _c64(np.asarray(vecs, dtype=np.float64).reshape(-1, 12))
        if self.vecs is not None:
            A = self.vecs.shape[0]
            dsp = np.ones(2)
            dof = np.zeros(2)
            rot = np.zeros(3)
            ang = np.zeros((3, A))
            dso = dsd = cor = 0.0
        else:
            dsp = np.asarray(geom.det_spacing, dtype=np.float64)
            dof = np.asarray(geom.det_offset, dtype=np.float64)
            rot = np.asarray(geom.det_rotation, dtype=np.float64)
            ang = np.asarray(geom.angles, dtype=np.float64)
            if ang.ndim == 1:
                ang = np.vstack([ang, np.zeros_like(ang), np.zeros_like(ang)])
            dso, dsd = float(geom.DSO), float(geom.DSD)
            cor = float(getattr(geom, "COR", 0.0))
        self.phi, self.theta, self.psi = _c64(ang[0]), _c64(ang[1]), _c64(ang[2])
        A = self.phi.size
        self.s = Cone3D(
            nx=int(vs[2]), ny=int(vs[1]), nz=int(vs[0]),
            sx=float(sp[2]), sy=float(sp[1]), sz=float(sp[0]),
            ox=float(of[2]), oy=float(of[1]), oz=float(of[0]),
            nu=int(ds[1]), nv=int(ds[0]), du=float(dsp[1]), dv=float(dsp[0]),
            off_u=float(dof[1]), off_v=float(dof[0]),
            pitch=float(rot[0]), roll=float(rot[1]), yaw=float(rot[2]),
            dso=dso, dsd=dsd, cor=cor,
            nangles=A, phi=_d(self.phi), theta=_d(self.theta), psi=_d(self.psi),
            vecs=_d(self.vecs) if self.vecs is not None else None,
        )
        self.shape_vol = (int(vs[0]), int(vs[1]), int(vs[2]))
        self.shape_sino = (A, int(ds[0]), int(ds[1]))
        self.A = A

    def plan_fwd(self):
        arrs = [np.zeros((self.A, 3)) for _ in range(4)]
        lib().orc_plan3d_fwd(C.byref(self.s), *[_d(x) for x in arrs])
        return arrs  # origin, du, dv, src

    def plan_bp(self):
        arrs = [np.zeros((self.A, 3)) for _ in range(4)]
        src = np.zeros(3)
        lib().orc_plan3d_bp(C.byref(self.s), *[_d(x) for x in arrs], _d(src))
        return arrs + [src]  # origin, dx, dy, dz, src

    def ray(self, a, iu, iv, plan=None):
        org, du, dv, src = plan if plan is not None else self.plan_fwd()
        s = np.zeros(3)
        d = np.zeros(3)
        t0 = np.zeros(1)
        dt = np.zeros(1)
        n = lib().orc_ray3d(C.byref(self.s), _d(org), _d(du), _d(dv), _d(src), a, iu, iv,
                            _d(s), _d(d), _d(t0), _d(dt))
        return int(n), s, d, float(t0[0]), float(dt[0])


def fwd3d(geom, vol, angles=None, count=False):
    c = Cone(geom)
    vol = _c64(vol)
    assert vol.shape == c.shape_vol
    a0, a1 = (0, c.A) if angles is None else angles
    sino = np.full(c.shape_sino, np.nan)
    ns = _i64(0)
    lib().orc_fwd3d(C.byref(c.s), _d(vol), _d(sino), a0, a1, C.byref(ns))
    return (sino, int(ns.value)) if count else sino


def adj3d(geom, sino, angles=None):
    c = Cone(geom)
    sino = _c64(sino)
    assert sino.shape == c.shape_sino
    a0, a1 = (0, c.A) if angles is None else angles
    vol = np.zeros(c.shape_vol)
    lib().orc_adj3d(C.byref(c.s), _d(sino), _d(vol), a0, a1)
    return vol


def bp3d(geom, sino, slabs=None, vol=None, matched=False, fdk=False):
    """FP64 3D backprojection: the reference's (default), geometry-matched, or matched with
    the FDK distance weight."""
    c = Cone(geom)
    if fdk and c.vecs is not None:
        raise ValueError("FDK weighting needs a circular orbit, not views")
    sino = _c64(sino)
    assert sino.shape == c.shape_sino
    out = np.zeros(c.shape_vol) if vol is None else _c64(vol).copy()
    z0, z1 = (0, c.shape_vol[0]) if slabs is None else slabs
    fn = lib().orc_bp3d_fdk if fdk else (lib().orc_bp3d_matched if matched else lib().orc_bp3d)
    fn(C.byref(c.s), _d(sino), _d(out), z0, z1)
    return out


def bp3d_T(geom, vol, angles=None, matched=False, fdk=False):
    """FP64 transpose of ``bp3d`` (same modes) applied to a volume, into a projection stack;
    ``angles`` = (a0, a1) restricts it to those projections (the others stay 0)."""
    c = Cone(geom)
    if fdk and c.vecs is not None:
        raise ValueError("FDK weighting needs a circular orbit, not views")
    vol = _c64(vol)
    assert vol.shape == c.shape_vol
    out = np.zeros(c.shape_sino)
    a0, a1 = (0, c.A) if angles is None else angles
    lib().orc_bp3d_T(C.byref(c.s), 2 if fdk else (1 if matched else 0), _d(vol), _d(out), a0, a1)
    return out


def fwd3d_f32(geom, vol, angles=None):
    """FP32 CPU port of the reference's 3D forward (timing only) over projections (a0, a1)."""
    c = Cone(geom)
    vol = np.ascontiguousarray(vol, dtype=np.float32)
    a0, a1 = (0, c.A) if angles is None else angles
    sino = np.zeros(c.shape_sino, dtype=np.float32)
    lib().orc_fwd3d_f32(C.byref(c.s), _f(vol), _f(sino), a0, a1)
    return sino


def bp3d_f32(geom, sino, slabs=None):
    """FP32 CPU port of the reference's 3D backprojection (timing only) over slices (z0, z1)."""
    c = Cone(geom)
    sino = np.ascontiguousarray(sino, dtype=np.float32)
    out = np.zeros(c.shape_vol, dtype=np.float32)
    z0, z1 = (0, c.shape_vol[0]) if slabs is None else slabs
    lib().orc_bp3d_f32(C.byref(c.s), _f(sino), _f(out), z0, z1)
    return out
