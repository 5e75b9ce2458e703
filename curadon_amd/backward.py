"""``backward`` (python/standalone/src/curadon/backward.py:7-29): voxel-driven
backprojection, accumulated (+=) into a caller-provided volume, which is returned.

Also accepted: ``backward(sinogram, geom)`` (the volume is then allocated with zeros) and
``backward(sinogram, geom, volume=...)``. ``mode`` selects the weighting / projection (see
``backward``).
"""
from __future__ import annotations

import torch

from . import _lib as _C
from .forward import _split
from .geometry import ConeGeometry, FanGeometry


def backward(sinogram, volume=None, geom=None, *, volume_out=None, mode="reference"):
    """``mode``: 'reference' (default, the reference's weighting and projection), 'fbp' (2D:
    squared magnification weight), 'matched' / 'fdk' (3D: the forward model's projection,
    unweighted / FDK-weighted)."""
    volume, geom = _split(volume, geom, None, volume_out, "volume")
    if isinstance(geom, FanGeometry):
        if mode not in ("reference", "fbp"):
            raise ValueError(f"fan-beam backward: unknown mode {mode!r}")
        if volume is None:
            volume = torch.zeros(tuple(sinogram.shape[:-2]) + tuple(int(v) for v in geom.vol_shape),
                                 dtype=_C.dtype_of_code(geom.plan.vol_prec),
                                 device=sinogram.device)
        _C.backward_2d(volume, sinogram, geom.plan, weighting=mode)
        return volume
    if isinstance(geom, ConeGeometry):
        if volume is None:
            volume = torch.zeros(tuple(sinogram.shape[:-3]) + tuple(int(v) for v in geom.vol_shape),
                                 dtype=_C.dtype_of_code(geom.plan.vol_prec),
                                 device=sinogram.device)
        _C.backward_3d_plan(volume, sinogram, geom.plan, mode=mode)
        return volume
    raise ValueError("Unknown geometry type")


def adjoint(sinogram, volume=None, geom=None, *, volume_out=None):
    """Matched adjoint A^T of ``forward`` (exact transpose), accumulated into ``volume``."""
    volume, geom = _split(volume, geom, None, volume_out, "volume")
    if isinstance(geom, FanGeometry):
        if volume is None:
            volume = torch.zeros(tuple(sinogram.shape[:-2]) + tuple(int(v) for v in geom.vol_shape),
                                 dtype=_C.dtype_of_code(geom.plan.vol_prec),
                                 device=sinogram.device)
        _C.adjoint_2d(volume, sinogram, geom.plan)
        return volume
    if isinstance(geom, ConeGeometry):
        if volume is None:
            volume = torch.zeros(tuple(sinogram.shape[:-3]) + tuple(int(v) for v in geom.vol_shape),
                                 dtype=_C.dtype_of_code(geom.plan.vol_prec),
                                 device=sinogram.device)
        _C.adjoint_3d_plan(volume, sinogram, geom.plan)
        return volume
    raise ValueError("Unknown geometry type")


def backward_transpose(volume, geom, sinogram=None, *, mode="reference"):
    """Transpose of ``backward``: ``sinogram := B^T volume`` (overwritten, or allocated when
    None) with the same ``mode`` (fan beams 'reference' or 'fbp'; cone beams 'reference',
    'matched' or 'fdk'); the gradient of the voxel-driven backprojection. Returns the
    sinogram."""
    if isinstance(geom, FanGeometry):
        if mode not in ("reference", "fbp"):
            raise ValueError(f"fan-beam backward_transpose: unknown mode {mode!r}")
        if sinogram is None:
            sinogram = torch.empty(tuple(volume.shape[:-2]) + tuple(geom.sinogram_shape()),
                                   dtype=_C.dtype_of_code(geom.plan.det_prec),
                                   device=volume.device)
        _C.backward_2d_transpose(volume, sinogram, geom.plan, weighting=mode)
        return sinogram
    if isinstance(geom, ConeGeometry):
        if sinogram is None:
            sinogram = torch.empty(tuple(volume.shape[:-3]) + tuple(geom.sinogram_shape()),
                                   dtype=_C.dtype_of_code(geom.plan.det_prec),
                                   device=volume.device)
        _C.backward_3d_transpose(volume, sinogram, geom.plan, mode=mode)
        return sinogram
    raise ValueError("Unknown geometry type")
