"""Autograd functions (python/standalone/src/curadon/diffable.py:14-38).

``TorchForward.apply(volume, geom)`` returns ``A volume``; its gradient is the matched
adjoint ``A^T grad`` (exact transpose of the ray march), not the voxel-driven backprojector
the reference pairs it with (diffable.py:22-24).

``TorchBackprojection.apply(sinogram, geom)`` returns the voxel-driven backprojection
``B sinogram``; its gradient is the exact transpose ``B^T grad`` (the backprojector's own
projection and weights), where the reference uses the forward projection ``A grad``
(diffable.py:35-38), an approximation since ``B`` is not the transpose of ``A``. For cone beams
with a voxel at or behind a source, where ``B^T`` is refused, the gradient falls back to
``A grad``.

Each function returns exactly one gradient per input.
"""
from __future__ import annotations

import torch
from torch.autograd import Function

from .backward import adjoint, backward, backward_transpose
from .forward import forward
from .geometry import FanGeometry


class TorchForward(Function):
    @staticmethod
    def forward(ctx, volume, geom):
        ctx.geom = geom
        return forward(volume.contiguous(), geom)

    @staticmethod
    def backward(ctx, grad_sino):
        return adjoint(grad_sino.contiguous(), ctx.geom), None


class TorchAdjoint(Function):
    """``A^T sinogram`` with gradient ``A grad`` (exact pair with TorchForward)."""

    @staticmethod
    def forward(ctx, sino, geom):
        ctx.geom = geom
        return adjoint(sino.contiguous(), geom)

    @staticmethod
    def backward(ctx, grad_vol):
        return forward(grad_vol.contiguous(), ctx.geom), None


class TorchBackprojection(Function):
    @staticmethod
    def forward(ctx, sino, geom):
        ctx.geom = geom
        return backward(sino.contiguous(), geom)

    @staticmethod
    def backward(ctx, grad_vol):
        if isinstance(ctx.geom, FanGeometry):
            return backward_transpose(grad_vol.contiguous(), ctx.geom), None
        try:
            return backward_transpose(grad_vol.contiguous(), ctx.geom), None
        except NotImplementedError:
            return forward(grad_vol.contiguous(), ctx.geom), None
