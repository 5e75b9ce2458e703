"""curadon_amd -- MI355X (gfx950) projection operators with curadon's Python API.

Drop-in names (python/standalone/src/curadon/__init__.py:1-12): FanGeometry, ConeGeometry,
forward, backward, TorchForward, TorchBackprojection, plan_2d, forward_2d, backward_2d,
forward_3d, backward_3d. Additions: adjoint / adjoint_2d / TorchAdjoint (matched transpose),
backward_transpose / backward_2d_transpose / backward_3d_transpose (transpose of the
backprojection, the gradient TorchBackprojection uses), fbp / fbp_filter (filtered
backprojection, FDK for cone beam). Stream and event pools under
curadon's names (event, event_view, event_pool, stream, stream_view, stream_pool,
get_next_stream, get_stream, get_next_event, get_event) come from curadon_amd.streams.
Usage: ``import curadon_amd as curadon``.
"""
from .geometry import ConeGeometry, FanGeometry
from . import _lib
from ._lib import (CuradonError, adjoint_2d, backward_2d, backward_2d_transpose, backward_3d,
                   backward_3d_transpose, forward_2d, forward_3d, plan_2d, plan_3d)
from .forward import forward
from .backward import adjoint, backward, backward_transpose
from .diffable import TorchAdjoint, TorchBackprojection, TorchForward
from .reconstruction import fbp, fbp_filter
from .streams import (event, event_pool, event_view, get_event, get_next_event, get_next_stream,
                      get_stream, stream, stream_pool, stream_view)

__all__ = [
    "FanGeometry", "ConeGeometry", "forward", "backward", "adjoint", "TorchForward",
    "TorchBackprojection", "TorchAdjoint", "plan_2d", "forward_2d", "backward_2d",
    "adjoint_2d", "forward_3d", "backward_3d", "plan_3d", "CuradonError", "fbp", "fbp_filter",
    "backward_transpose", "backward_2d_transpose", "backward_3d_transpose",
    "event", "event_view", "event_pool", "stream", "stream_view", "stream_pool",
    "get_next_stream", "get_stream", "get_next_event", "get_event",
]
