"""``forward`` (python/standalone/src/curadon/forward.py:7-28): ray-driven forward projection
into a caller-provided sinogram, which is returned.

Also accepted, for the reference's older two-argument callers (python/tests/
test_forward_2d.py:82, diffable.py:17): ``forward(volume, geom)`` and
``forward(volume, geom, sinogram=...)``; the sinogram is then allocated (or taken from the
keyword). Work is enqueued on the current torch stream of the volume's device.
"""
from __future__ import annotations

import torch

from . import _lib as _C
from .geometry import ConeGeometry, FanGeometry

_GEOMS = (FanGeometry, ConeGeometry)


def _split(first, second, geom_kw, out_kw, out_name):
    """Sort (output, geometry) out of the positional / keyword forms."""
    out, geom = None, geom_kw
    for a in (first, second):
        if isinstance(a, _GEOMS):
            if geom is not None and geom is not a:
                raise TypeError("geometry given twice")
            geom = a
        elif a is not None:
            if out is not None:
                raise TypeError(f"{out_name} given twice")
            out = a
    if out_kw is not None:
        if out is not None and out is not out_kw:
            raise TypeError(f"{out_name} given twice")
        out = out_kw
    return out, geom


def forward(volume, sinogram=None, geom=None, *, sinogram_out=None):
    sinogram, geom = _split(sinogram, geom, None, sinogram_out, "sinogram")
    if isinstance(geom, FanGeometry):
        if sinogram is None:
            sinogram = torch.empty(tuple(volume.shape[:-2]) + geom.sinogram_shape(),
                                   dtype=_C.dtype_of_code(geom.plan.det_prec),
                                   device=volume.device)
        _C.forward_2d(volume, sinogram, geom.plan)
        return sinogram
    if isinstance(geom, ConeGeometry):
        if sinogram is None:
            sinogram = torch.empty(tuple(volume.shape[:-3]) + geom.sinogram_shape(),
                                   dtype=_C.dtype_of_code(geom.plan.det_prec),
                                   device=volume.device)
        _C.forward_3d_plan(volume, sinogram, geom.plan)
        return sinogram
    raise ValueError("Unknown geometry type")
