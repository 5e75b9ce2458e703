"""cmi_forward_2d_noalloc (the 2D forward with its staging scratch in the library's static
device array) against cmi_forward_2d (stream-ordered scratch), through the C-ABI: the same
kernels, so the sinograms must be bit-identical. Two calls in a row, the second with a smaller
input, so that the second one reads a scratch array that still holds the first call's larger
copies (every element it reads must have been rewritten). Sizes beyond the static maximum are
refused before anything is enqueued."""
import numpy as np
import pytest
import torch

import curadon_amd as cm
from curadon_amd import _lib, phantom

pytestmark = pytest.mark.gpu


def _geom(n, ny, D, A, **kw):
    return cm.FanGeometry(DSD=3.0 * n, DSO=2.0 * n, angles=np.linspace(0, 2 * np.pi, A, endpoint=False),
                          vol_shape=(ny, n), det_shape=D, **kw)


def _both(geom, x):
    s_ref = torch.full((*x.shape[:-2], *geom.sinogram_shape()), float("nan"), device=x.device,
                       dtype=_lib.dtype_of_code(geom.plan.det_prec))
    s_new = s_ref.clone()
    _lib.forward_2d(x, s_ref, geom.plan)
    _lib.forward_2d_noalloc(x, s_new, geom.plan)
    torch.cuda.synchronize()
    return s_ref, s_new


def test_noalloc_equals_forward_two_calls(gpu):
    g_big = _geom(512, 512, 724, 256)
    x_big = torch.stack([torch.tensor(phantom.ellipses_2d(512, seed=s), dtype=torch.float32)
                         for s in range(2)]).to(gpu)
    r, n = _both(g_big, x_big)
    assert torch.equal(r, n)
    assert float(r.abs().max()) > 0
    # smaller, non-square input: the static array still holds the 512^2 copies
    g_small = _geom(96, 80, 131, 60)
    x_small = torch.rand(80, 96, device=gpu, generator=torch.Generator(device=gpu).manual_seed(3))
    r, n = _both(g_small, x_small)
    assert torch.equal(r, n)
    assert not torch.isnan(n).any()


def test_noalloc_half_storage(gpu):
    g = _geom(128, 128, 182, 90, vol_prec=torch.float16, sino_prec=torch.float16)
    x = torch.rand(128, 128, device=gpu, generator=torch.Generator(device=gpu).manual_seed(4))
    r, n = _both(g, x.half())
    assert torch.equal(r, n)


def test_noalloc_refuses_oversized_call(gpu):
    # 2 (2900 + 4)^2 4 bytes = 67.4 MB > 64 MiB: refused on the host, nothing enqueued
    g = _geom(2900, 2900, 8, 2)
    x = torch.zeros(2900, 2900, device=gpu)
    s = torch.full((2, 8), 5.0, device=gpu)
    with pytest.raises(ValueError, match="static maximum"):
        _lib.forward_2d_noalloc(x, s, g.plan)
    torch.cuda.synchronize()
    assert torch.all(s == 5.0)
    # the largest square FP32 image that fits still runs
    n = int((_lib.NOALLOC_SCRATCH_BYTES / 8) ** 0.5) - 4
    g2 = _geom(n, n, 8, 2)
    x2 = torch.rand(n, n, device=gpu)
    r, m = _both(g2, x2)
    assert torch.equal(r, m)
