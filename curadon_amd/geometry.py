"""Geometry descriptors with the constructor signatures, attributes and validation errors of
curadon's ``FanGeometry`` / ``ConeGeometry`` (python/standalone/src/curadon/geometry.py:53-336).

Differences from the reference, all deliberate:
  * a scalar ``DSO`` is broadcast from ``DSO`` (the reference broadcasts ``DSD`` into it,
    geometry.py:165, which puts the detector through the isocentre);
  * the native plan is built on first use (``geom.plan``) instead of in ``__init__``, so a
    geometry can be described without a GPU;
  * ``vol_prec`` / ``sino_prec`` also accept torch dtypes and 'bf16';
  * ``ConeGeometry`` owns a cached native 3D plan (the reference rebuilds its geometry tables
    on every call).
"""
from __future__ import annotations

from typing import Any, List, Optional, Tuple, Union

import numpy as np


def _as_float(val: Any) -> float:
    """float(val); a one-element array goes through .item() first (float() of an array with
    ndim > 0 is deprecated in NumPy)."""
    if isinstance(val, np.ndarray) and val.size == 1:
        val = val.item()
    return float(val)


def _rot_yzy(phi, theta, psi):
    """The trajectory rotation of the reference's cone beam, the matrix written in
    include/curadon/detail/rotation.hpp:18-21."""
    cf, sf, ct, st, cp, sp = (np.cos(phi), np.sin(phi), np.cos(theta), np.sin(theta),
                              np.cos(psi), np.sin(psi))
    return np.array([[cf * cp * ct - sf * sp, -st * cf, sf * cp + sp * cf * ct],
                     [st * cp, ct, sp * st],
                     [-sf * cp * ct - sp * cf, sf * st, -sf * sp * ct + cp * cf]])


def _rot_rpy(roll, pitch, yaw):
    """The detector rotation of the reference's cone beam (rotation.hpp:31-48)."""
    cr, sr, cp, sp, cy, sy = (np.cos(roll), np.sin(roll), np.cos(pitch), np.sin(pitch),
                              np.cos(yaw), np.sin(yaw))
    return np.array([[cr * cp, cr * sp * sy - sr * cy, cr * sp * cy + sr * sy],
                     [sr * cp, sr * sp * sy + cr * cy, sr * sp * cy - cr * sy],
                     [-sp, cp * sy, cp * cy]])


def _fp(a) -> bytes:
    """Bytes of an array-like as FP64, for plan fingerprints."""
    a = np.asarray(a, dtype=np.float64)
    return bytes(repr(a.shape), "ascii") + a.tobytes()


def is_convertible_to_float(val: Any):
    try:
        _as_float(val)
        return True
    except (ValueError, TypeError):
        return False


def geom_to_str(geom):
    p = ["-----", "Geometry parameters",
         f"Distance from source to detector (DSD) = {geom.DSD}",
         f"Distance from source to origin (DSO) = {geom.DSO}"]
    if isinstance(geom, FanGeometry):
        p.append(f"Projections angles = {np.rad2deg(geom.angles)}")
    p += ["-----", "Detector parameters",
          f"Number of pixels (det_shape) = {geom.det_shape}",
          f"Size of each pixel (det_spacing) = {geom.det_spacing}",
          f"Total size of the detector (det_size) = {np.asarray(geom.det_shape) * geom.det_spacing}",
          f"Detector rotation (det_rotation) = {geom.det_rotation}",
          f"Offset of detector (det_offset) = {geom.det_offset}",
          "-----", "Image parameters",
          f"Number of voxels (vol_shape) = {geom.vol_shape}",
          f"Size of each voxel (vol_spacing) = {geom.vol_spacing}",
          f"Total size of the image (vol_size) = {geom.vol_shape * geom.vol_spacing}",
          f"Offset of image from origin (vol_offset) = {geom.vol_offset}"]
    return "\n".join(p)


class FanGeometry:
    """2D fan-beam geometry with a flat detector (geometry.py:136-235)."""

    def to_vec(self):
        """ASTRA-like vectors ``[src.x, src.y, det.x, det.y, u.x, u.y]`` per angle
        (geometry.py:54-79): source at (det_offset, DSO), detector centre at (0, -(DSD - DSO))
        and cell step (det_spacing, 0), rotated by ``[[c, s], [-s, c]]``. With det_offset = 0
        and det_spacing = 1 these are the reference's vectors; otherwise the source shift that
        the reference's forward applies (plan_2d.hpp:76-80) and the cell spacing are kept, so
        that ``from_vec(to_vec())`` describes the same projector. A geometry made by
        ``from_vec`` returns its views."""
        if getattr(self, "vecs", None) is not None:
            return self.vecs.tolist()
        out = []
        if getattr(self, "exact_geometry", False):
            # the views of cmi_plan2d_create_ex with exact geometry
            p = self.det_rotation
            for alpha, dsd, dso in zip(self.angles, self.DSD, self.DSO):
                c, s = np.cos(alpha), np.sin(alpha)
                R = np.array([[c, s], [-s, c]])
                src = R @ np.array([-self.COR, dso])
                det = R @ np.array([self.det_offset - self.COR, -(dsd - dso)])
                u = R @ (self.det_spacing * np.array([np.cos(p), np.sin(p)]))
                out.append([*src.tolist(), *det.tolist(), *u.tolist()])
            return out
        for alpha, dsd, dso in zip(self.angles, self.DSD, self.DSO):
            c, s = np.cos(alpha), np.sin(alpha)
            R = np.array([[c, s], [-s, c]])
            src = R @ np.array([self.det_offset, dso])
            det = R @ np.array([0.0, -(dsd - dso)])
            u = R @ np.array([self.det_spacing, 0.0])
            out.append([*src.tolist(), *det.tolist(), *u.tolist()])
        return out

    @staticmethod
    def from_vec(vol_shape: Tuple[int, int], det_shape: int, vecs: List[List[float]], **kwargs):
        """Build a FanGeometry from ASTRA-like vectors ``[src.x, src.y, det.x, det.y, u.x,
        u.y]`` per view (world coordinates; ``det`` the detector centre, ``u`` the step from
        cell 0's centre to cell 1's). The geometry keeps the views (``vecs``) and its plan
        projects through them exactly (cmi_plan2d_create_vec): a moved source, a shifted or
        tilted detector and a per-view cell spacing are all honoured. The reference's version
        (geometry.py:84-133) keeps only the summary below and rejects unequal spacings; the
        summary is still set, for display and for code that reads it: DSD = |src - det|,
        DSO = |src|, angle = atan2(src - det) mod 2 pi, det_spacing = |u| of the first view."""
        vecs = np.asarray(vecs, dtype=np.float64)
        if vecs.ndim != 2 or vecs.shape[1] != 6:
            raise AttributeError(f"vecs must have shape (nangles, 6), got {vecs.shape}")
        det_spacing = float(np.linalg.norm(vecs[0][-2:]))
        dsd, dso, angles = [], [], []
        for v in vecs:
            sx, sy, dx, dy, ux, uy = v
            src = np.array([sx, sy])
            s2d = src - np.array([dx, dy])
            dsd.append(float(np.linalg.norm(s2d)))
            dso.append(float(np.linalg.norm(src)))
            angles.append(float(np.arctan2(s2d[0], s2d[1]) % (2 * np.pi)))
            if not np.linalg.norm([ux, uy]) > 0.0:
                raise AttributeError("Every view needs a non-zero cell step u")
        g = FanGeometry(DSD=dsd, DSO=dso, angles=angles, vol_shape=vol_shape,
                        det_shape=det_shape, det_spacing=det_spacing, **kwargs)
        g.vecs = np.array(vecs)
        return g

    def __init__(self, DSD: Union[float, List[float]], DSO: Union[float, List[float]],
                 angles: List[float], vol_shape: Tuple[int, int], det_shape: int,
                 det_spacing: Optional[float] = None, det_offset: Optional[float] = None,
                 det_rotation: Optional[float] = None,
                 vol_spacing: Optional[Tuple[float, float]] = None,
                 vol_offset: Optional[Tuple[float, float]] = None, device: int = 0,
                 vol_prec=32, sino_prec=32, *, COR: Optional[float] = None,
                 exact_geometry: bool = False):
        """``COR`` and ``exact_geometry`` are additions (keyword only). With
        ``exact_geometry=True`` the detector geometry is honoured consistently by the forward,
        backward and adjoint: det_offset shifts the detector along its axis, det_rotation
        turns it about its centre, and COR is the distance along the detector axis from the
        line through the source and the detector centre to the centre of rotation. Without it
        (default) the reference's semantics hold: the forward moves the source by det_offset,
        the backprojection ignores it, and det_rotation and COR are stored and unused
        (plan_2d.hpp:76-80, 121-122)."""
        self.angles = np.asarray(angles, dtype=np.float64)
        if self.angles.ndim != 1:
            raise AttributeError(f"Angles must be 1D array, got {str(self.angles.shape)}")
        self.nangles = self.angles.shape[0]

        if is_convertible_to_float(DSD) is True:
            self.DSD = np.full(self.nangles, _as_float(DSD))
        else:
            self.DSD = np.asarray(DSD, dtype=np.float64)
            if self.DSD.ndim != 1 or self.DSD.size != self.nangles:
                raise AttributeError(
                    f"DSD must be a single float, or a list of floats nangles entries "
                    f"(is {self.DSD.size}, expected {self.nangles})")

        if is_convertible_to_float(DSO) is True:
            self.DSO = np.full(self.nangles, _as_float(DSO))
        else:
            self.DSO = np.asarray(DSO, dtype=np.float64)
            if self.DSO.ndim != 1 or self.DSO.size != self.nangles:
                raise AttributeError(
                    f"DSO must be a single float, or a list of floats nangles entries "
                    f"(is {self.DSO.size}, expected {self.nangles})")

        self.vol_shape = np.asarray(vol_shape)
        if self.vol_shape.shape != (2,):
            raise AttributeError(f"vol_shape must be two values, got {str(self.vol_shape.shape)}")

        if is_convertible_to_float(det_shape) is False:
            raise AttributeError(f"det_shape must to be float, got {type(det_shape)}")
        self.det_shape = int(_as_float(det_shape))

        if det_spacing is not None:
            if is_convertible_to_float(det_spacing) is False:
                raise TypeError(f"Expected det_spacing to be float, got {type(det_spacing)}")
            self.det_spacing = _as_float(det_spacing)
        else:
            self.det_spacing = 1.0

        if det_offset is not None:
            if is_convertible_to_float(det_offset) is False:
                raise TypeError(f"Expected det_offset to be float, got {type(det_offset)}")
            self.det_offset = _as_float(det_offset)
        else:
            self.det_offset = 0.0

        if det_rotation is not None:
            if is_convertible_to_float(det_rotation) is False:
                raise TypeError(f"Expected det_rotation to be float, got {type(det_rotation)}")
            self.det_rotation = _as_float(det_rotation)
        else:
            self.det_rotation = 0.0

        if vol_spacing is not None:
            self.vol_spacing = np.asarray(vol_spacing, dtype=np.float64)
            if self.vol_spacing.shape != (2,):
                raise AttributeError(
                    f"vol_spacing must be two values, got {str(self.vol_spacing.shape)}")
        else:
            self.vol_spacing = np.ones(self.vol_shape.shape)

        if vol_offset is not None:
            self.vol_offset = np.asarray(vol_offset, dtype=np.float64)
            if self.vol_offset.shape != (2,):
                raise AttributeError(
                    f"vol_offset must be two values, got {str(self.vol_offset.shape)}")
        else:
            self.vol_offset = np.zeros(self.vol_shape.shape)

        if COR is not None:
            if is_convertible_to_float(COR) is False:
                raise TypeError(f"Expected COR to be float, got {type(COR)}")
            self.COR = _as_float(COR)
        else:
            self.COR = 0.0
        self.exact_geometry = bool(exact_geometry)
        self.device = int(device)
        self.vol_prec = vol_prec
        self.sino_prec = sino_prec
        self.vecs = None  # ASTRA-like views when built by from_vec
        self._plan = None
        self._plan_key = None

    def fingerprint(self):
        """Every attribute the native plan is built from, as a hashable tuple: the plan is
        rebuilt when any of them changes after it was made."""
        return (_fp(self.angles), _fp(self.DSD), _fp(self.DSO), _fp(self.vol_shape),
                _fp(self.vol_spacing), _fp(self.vol_offset), int(self.det_shape),
                _as_float(self.det_spacing), _as_float(self.det_offset),
                _as_float(self.det_rotation), _as_float(self.COR), bool(self.exact_geometry),
                int(self.device), str(self.vol_prec), str(self.sino_prec),
                None if self.vecs is None else _fp(self.vecs))

    @property
    def plan(self):
        """The native plan (``curadon_amd._lib.plan_2d``), built on first access and rebuilt
        whenever an attribute it depends on has changed since (``fingerprint``): from the views
        for a geometry made by ``from_vec``, else from angles / DSO / DSD, det_rotation and COR
        (the reference passes COR 0, geometry.py:226-228, and its plan ignores both), with the
        exact-geometry flag when ``exact_geometry`` is set."""
        key = self.fingerprint()
        if self._plan is None or self._plan_key != key:
            from . import _lib
            self._plan = None
            if self.vecs is not None:
                self._plan = _lib.plan_2d.from_views(
                    self.device, self.vol_prec, self.vol_shape, self.vol_spacing,
                    self.vol_offset, self.sino_prec, self.det_shape, self.vecs)
            else:
                self._plan = _lib.plan_2d(self.device, self.vol_prec, self.vol_shape,
                                          self.vol_spacing, self.vol_offset, self.sino_prec,
                                          self.det_shape, self.det_spacing, self.det_offset,
                                          self.DSO, self.DSD, self.angles, self.det_rotation,
                                          self.COR, exact_geometry=self.exact_geometry)
            self._plan_key = key
        return self._plan

    def is_centred_circular(self):
        """True when the projector is the reference's centred circular fan beam with the
        detector on its central ray: no views, and (with exact geometry) no detector offset,
        rotation or COR. FBP weighting and the serving tables assume this."""
        if self.vecs is not None:
            return False
        if self.exact_geometry:
            return self.det_offset == 0.0 and self.det_rotation == 0.0 and self.COR == 0.0
        return True

    def sinogram_shape(self):
        return (self.nangles, self.det_shape)

    def __str__(self):
        return geom_to_str(self)


class ConeGeometry:
    """3D cone-beam geometry with a flat panel (geometry.py:237-336). Axis order as the
    reference's Python side: vol_shape (z, y, x), det_shape (v, u); angles (A,) or (3, A) as
    (phi, theta, psi); det_rotation (pitch, roll, yaw)."""

    def __init__(self, DSD: float, DSO: float, angles, vol_shape: Tuple[int, int, int],
                 det_shape: Tuple[int, int], det_spacing=None, det_offset=None,
                 det_rotation=None, vol_spacing=None, vol_offset=None, COR=None, device: int = 0,
                 vol_prec=32, sino_prec=32):
        if is_convertible_to_float(DSD) is False:
            raise TypeError(f"Expected DSD to be float, got {type(DSD)}")
        self.DSD = _as_float(DSD)
        if is_convertible_to_float(DSO) is False:
            raise TypeError(f"Expected DSO to be float, got {type(DSO)}")
        self.DSO = _as_float(DSO)

        self.det_shape = np.asarray(det_shape)
        if self.det_shape.shape != (2,):
            raise AttributeError(
                f"Detector shape must be two values, got {str(self.det_shape.shape)}")
        self.vol_shape = np.asarray(vol_shape)
        if self.vol_shape.shape != (3,):
            raise AttributeError(
                f"Volume shape must be three values, got {str(self.vol_shape.shape)}")

        self.angles = np.asarray(angles, dtype=np.float64)
        if self.angles.ndim == 1:
            self.nangles = self.angles.shape[0]
            z = np.zeros((self.nangles,), dtype=np.float64)
            self.angles = np.vstack((self.angles, z, z))
        elif self.angles.ndim == 2:
            if self.angles.shape[0] != 3:
                raise AttributeError(
                    f"Angles must either be a 1D array, or a 2D array of shape (3, nangles), "
                    f"got {str(self.angles.shape)}")
            self.nangles = self.angles.shape[1]
        else:
            raise AttributeError(
                f"Angles must either be a 1D array, or a 2D array of shape (3, nangles), "
                f"got {str(self.angles.shape)}")

        def vec(val, n, name, default):
            if val is None:
                return default
            a = np.asarray(val, dtype=np.float64)
            if a.shape != (n,):
                raise AttributeError(f"{name} must be {'two' if n == 2 else 'three'} values, "
                                     f"got {str(a.shape)}")
            return a

        self.det_spacing = vec(det_spacing, 2, "det_spacing", np.ones(2))
        self.det_offset = vec(det_offset, 2, "det_offset", np.zeros(2))
        self.det_rotation = vec(det_rotation, 3, "det_rotation", np.zeros(3))
        if COR is not None:
            if is_convertible_to_float(COR) is False:
                raise TypeError(f"Expected COR to be float, got {type(COR)}")
            self.COR = _as_float(COR)
        else:
            self.COR = 0.0
        self.vol_spacing = vec(vol_spacing, 3, "vol_spacing", np.ones(3))
        self.vol_size = self.vol_shape * self.vol_spacing
        self.vol_offset = vec(vol_offset, 3, "vol_offset", np.zeros(3))
        self.device = int(device)
        self.vol_prec = vol_prec
        self.sino_prec = sino_prec
        self.vecs = None  # ASTRA-like views when built by from_vec
        self._plan = None
        self._plan_key = None

    def to_vec(self):
        """ASTRA-like cone views, one row per projection: (source x, y, z, panel centre x, y,
        z, u-step x, y, z, v-step x, y, z) in world coordinates (the volume's centre at
        vol_offset), the steps going from pixel (iu, iv) to (iu + 1, iv) and to (iu, iv + 1).
        For a geometry from Euler angles these are the points of the reference's forward set-up
        (forward_3d.hpp:118-190: detector roll / pitch / yaw, the shift by DSD - DSO, the
        trajectory rotation and the COR shift) before its translation into voxel coordinates,
        so ``from_vec(to_vec())`` projects like this geometry. The reference has no vector form
        for cone beam (its README.md:58-59 lists projection matrices as future work). A geometry
        made by ``from_vec`` returns its views."""
        if self.vecs is not None:
            return np.array(self.vecs)
        du, dv = float(self.det_spacing[1]), float(self.det_spacing[0])
        V, U = (int(v) for v in self.det_shape)
        pitch, roll, yaw = (float(v) for v in self.det_rotation)
        dod = self.DSD - self.DSO
        p00 = np.array([-du * (U / 2.0) + du * 0.5, -dv * (V / 2.0) + dv * 0.5, 0.0])
        pts = [p00, p00 + [du, 0.0, 0.0], p00 + [0.0, dv, 0.0]]
        Rd = _rot_rpy(roll, pitch, yaw)
        pts = [Rd @ p + [0.0, 0.0, dod] for p in pts]
        src0 = np.array([0.0, 0.0, -self.DSO])
        out = np.zeros((self.nangles, 12))
        for a in range(self.nangles):
            phi, theta, psi = (float(v) for v in self.angles[:, a])
            R = _rot_yzy(phi, theta, psi)
            cor = self.COR * np.array([np.sin(phi), 0.0, np.cos(phi)])
            q = [R @ p + cor for p in pts]
            s = R @ src0 + cor
            u, v = q[1] - q[0], q[2] - q[0]
            c = q[0] + (U - 1) / 2.0 * u + (V - 1) / 2.0 * v
            out[a] = np.concatenate([s, c, u, v])
        return out

    @staticmethod
    def from_vec(vol_shape, det_shape, vecs, **kwargs):
        """Build a ConeGeometry from ASTRA-like cone views (see ``to_vec``): any trajectory,
        panel position, tilt and pixel pitch per projection. The geometry keeps the views
        (``vecs``) and its plan projects through them (cmi_plan3d_create_vec): forward,
        adjoint and backward in 'matched' mode follow every view; 'reference' mode projects
        like 'matched' (the reference's backprojection frame exists for Euler angles only) and
        'fdk' is refused. A summary is set for display and for code that reads it: DSD =
        |source - panel centre|, DSO = |source|, angles = (atan2(-s.x, -s.z), 0, 0), det_spacing
        = (|v|, |u|) of the first view."""
        vecs = np.asarray(vecs, dtype=np.float64)
        if vecs.ndim != 2 or vecs.shape[1] != 12:
            raise AttributeError(f"vecs must have shape (nangles, 12), got {vecs.shape}")
        src, det = vecs[:, 0:3], vecs[:, 3:6]
        dsd = float(np.linalg.norm(src[0] - det[0]))
        dso = float(np.linalg.norm(src[0]))
        phi = np.arctan2(-src[:, 0], -src[:, 2]) % (2 * np.pi)
        spacing = (float(np.linalg.norm(vecs[0, 9:12])), float(np.linalg.norm(vecs[0, 6:9])))
        kw = dict(kwargs)
        kw.setdefault("det_spacing", spacing)
        g = ConeGeometry(DSD=dsd, DSO=dso, angles=phi, vol_shape=vol_shape, det_shape=det_shape,
                         **kw)
        g.vecs = np.array(vecs)
        return g

    def fingerprint(self):
        """Every attribute the native plan is built from, as a hashable tuple: the plan is
        rebuilt when any of them changes after it was made (the reference reads them on every
        call, python/standalone/src/curadon/forward.py:18-26)."""
        return (_fp(self.angles), _as_float(self.DSD), _as_float(self.DSO), _fp(self.vol_shape),
                _fp(self.vol_spacing), _fp(self.vol_offset), _fp(self.det_shape),
                _fp(self.det_spacing), _fp(self.det_offset), _fp(self.det_rotation),
                _as_float(self.COR), int(self.device), str(self.vol_prec), str(self.sino_prec),
                None if self.vecs is None else _fp(self.vecs))

    def _plan_args(self):
        return (self.vol_shape, self.vol_spacing, self.vol_offset, self.angles[0],
                self.angles[1], self.angles[2], self.det_shape, self.det_spacing,
                self.det_offset, self.det_rotation, self.DSO, self.DSD, self.COR)

    @property
    def plan(self):
        """The native 3D plan, built on first access and reused by every call until an
        attribute it depends on changes (``fingerprint``)."""
        key = self.fingerprint()
        if self._plan is None or self._plan_key != key:
            from . import _lib
            self._plan = None
            if self.vecs is not None:
                self._plan = _lib.plan_3d.from_views(
                    self.device, self.vol_prec, self.sino_prec, self.vol_shape,
                    self.vol_spacing, self.vol_offset, self.det_shape, self.vecs)
            else:
                self._plan = _lib.plan_3d(self.device, self.vol_prec, self.sino_prec,
                                          *self._plan_args())
            self._plan_key = key
        return self._plan

    def bp_matrices(self, mode="fdk", with_layout=False):
        """(A, 3, 4) FP32 backprojection matrices of ``mode`` for the current attributes, made
        on the host without building a device plan (cmi_cone3d_bp_matrices); with_layout also
        returns whether their U and W rows are free of y."""
        from . import _lib
        if self.vecs is not None:
            raise NotImplementedError("backprojection matrices of a geometry from views")
        return _lib.cone3d_bp_matrices(*self._plan_args(), mode=mode, with_layout=with_layout)

    def sinogram_shape(self):
        return (self.nangles, *[int(v) for v in self.det_shape])

    def __str__(self):
        return geom_to_str(self)
