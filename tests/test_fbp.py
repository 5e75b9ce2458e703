"""Filtered backprojection pipeline (examples/fbp.py:27-99) on the CPU: the filter, weights
and scale constants of curadon_amd.reconstruction are checked by reconstructing uniform objects with the
FP64 oracle's weighted backprojectors (a unit-density disc / ball must come back as 1), and the
windows / reference filtering against their definitions."""
import math

import numpy as np
import pytest
import torch

from curadon_amd import reconstruction as F
from curadon_amd.geometry import ConeGeometry, FanGeometry
from geoms import cone, fan


def fan_geom(n, D, A, dso, dsd):
    ang = np.linspace(0, 2 * np.pi, A, endpoint=False)
    return FanGeometry(DSD=dsd, DSO=dso, angles=ang, vol_shape=(n, n), det_shape=D)


@pytest.mark.parametrize("dso,dsd,D", [(512.0, 768.0, 400), (400.0, 700.0, 600)])
def test_fan_fbp_of_disc_is_unit_density(oracle, dso, dsd, D):
    n, A = 256, 512
    g = fan_geom(n, D, A, dso, dsd)
    c = np.arange(n) - n / 2 + 0.5
    X, Y = np.meshgrid(c, c)
    img = ((X ** 2 + Y ** 2) <= (0.3 * n) ** 2).astype(float)
    sino = oracle.fwd2d(g, img)
    q = F.fbp_filter(torch.tensor(sino), g).double().numpy()
    rec = oracle.bp2d(g, q, wpow=2)
    inner = (X ** 2 + Y ** 2) <= (0.25 * n) ** 2
    outer = (X ** 2 + Y ** 2) >= (0.36 * n) ** 2
    assert rec[inner].mean() == pytest.approx(1.0, abs=1e-2)
    assert abs(rec[outer & ((X ** 2 + Y ** 2) <= (0.45 * n) ** 2)].mean()) < 1e-2
    assert rec[inner].std() < 0.02


def test_cone_fdk_of_ball_is_unit_density(oracle):
    n = 48
    g = cone(n, 72, 72, 96, 96.0, 144.0)
    geom = ConeGeometry(DSD=g.DSD, DSO=g.DSO, angles=g.angles[0], vol_shape=g.vol_shape,
                        det_shape=g.det_shape)
    c = np.arange(n) - n / 2 + 0.5
    Z, Y, X = np.meshgrid(c, c, c, indexing="ij")
    r2 = X ** 2 + Y ** 2 + Z ** 2
    vol = (r2 <= (0.3 * n) ** 2).astype(float)
    sino = oracle.fwd3d(g, vol)
    # the forward is an unscaled sum of samples: divide by the step length to get line integrals
    c3 = oracle.Cone(g)
    pl = c3.plan_fwd()
    step = np.zeros(sino.shape)
    for a in range(0, 96, 8):
        for iv in range(72):
            for iu in range(72):
                nst, s, d, t0, dt = c3.ray(a, iu, iv, pl)
                step[a, iv, iu] = dt * np.linalg.norm(d) if nst >= 0 else 1.0
    # steps differ little between angles of a circular scan: reuse the sampled ones
    for a in range(96):
        step[a] = step[(a // 8) * 8]
    q = F.fbp_filter(torch.tensor(sino * step), geom).double().numpy()
    rec = oracle.bp3d(g, q, fdk=True)
    inner = r2 <= (0.2 * n) ** 2
    assert rec[inner].mean() == pytest.approx(1.0, abs=3e-2)


def test_windows_match_definitions():
    W = torch.linspace(-1, 1, 101, dtype=torch.float64)
    a = W.abs()
    torch.testing.assert_close(F._window(W, "ram-lak", 1.0), torch.ones_like(W))
    torch.testing.assert_close(F._window(W, "hann", 1.0), torch.cos(a * math.pi / 2) ** 2)
    torch.testing.assert_close(F._window(W, "hamming", 1.0), 0.54 + 0.46 * torch.cos(a * math.pi))
    sl = F._window(W, "shepp-logan", 1.0)
    torch.testing.assert_close(sl, torch.sin(math.pi * W / 2) / (math.pi * W / 2 + (W == 0)) * (W != 0) + (W == 0))
    half = F._window(W, "ram-lak", 0.5)
    assert (half[a > 0.5] == 0).all() and (half[a <= 0.5] == 1).all()
    with pytest.raises(RuntimeError):
        F._window(W, "nope", 1.0)


def test_ramp_response_is_band_limited_ramp():
    H = F.ramp_response(100).double()
    L = 2 * (H.shape[0] - 1)
    f = torch.arange(H.shape[0], dtype=torch.float64) / L
    # the band-limited ramp's response is |f| (cycles per cell) up to small truncation ripple
    assert torch.allclose(H, f, atol=2e-3)


def test_filter_rows_is_linear_convolution():
    r = np.random.default_rng(0)
    p = r.random((3, 37))
    H = F.ramp_response(37, window=None)
    got = F.filter_rows(torch.tensor(p), H).double().numpy()
    L = 2 * (H.shape[0] - 1)
    k = np.arange(-36, 37)
    h = np.zeros(k.size)
    h[k == 0] = 0.25
    h[k % 2 != 0] = -1.0 / (np.pi * k[k % 2 != 0]) ** 2
    expect = np.stack([np.convolve(row, h, mode="full")[36:36 + 37] for row in p])
    np.testing.assert_allclose(got, expect, atol=1e-6)


def test_toeplitz_filter_is_linear_convolution():
    """The Toeplitz matrix of the GPU filter path against numpy's linear convolution with the
    closed-form unwindowed kernel (h[0] = 1/4, h[odd k] = -1/(pi k)^2), and against filter_rows
    in FP64 for windowed filters."""
    r = np.random.default_rng(1)
    p = r.random((3, 37))
    got = (torch.tensor(p, dtype=torch.float32) @ F._toeplitz(37, None, 1.0, torch.device("cpu")))
    k = np.arange(-36, 37)
    h = np.zeros(k.size)
    h[k == 0] = 0.25
    h[k % 2 != 0] = -1.0 / (np.pi * k[k % 2 != 0]) ** 2
    expect = np.stack([np.convolve(row, h, mode="full")[36:36 + 37] for row in p])
    np.testing.assert_allclose(got.double().numpy(), expect, atol=1e-6)
    for D, win, s in ((64, "hann", 0.9), (181, "shepp-logan", 1.0), (100, "cosine", 0.5)):
        x = torch.tensor(r.random((5, D)))
        y = x.float() @ F._toeplitz(D, win, s, torch.device("cpu"))
        ref = F.filter_rows(x, F.ramp_response(D, win, s).double())
        assert ((y.double() - ref).norm() / ref.norm()).item() < 1e-6


def test_reference_filter_matches_examples_fbp():
    r = np.random.default_rng(1)
    s = torch.tensor(r.random((6, 41)), dtype=torch.float32)
    filt = torch.abs(torch.linspace(-1, 1, 41))
    expect = torch.real(torch.fft.ifft(torch.fft.fft(s, dim=1) * torch.fft.fftshift(filt), dim=1))
    torch.testing.assert_close(F.reference_filter(s), expect * (math.pi / 12), rtol=1e-5,
                               atol=1e-6)


def test_ramp_response_is_cached_and_copied():
    a = F.ramp_response(73, "hann", 0.8)
    b = F.ramp_response(73, "hann", 0.8)
    assert torch.equal(a, b) and a.data_ptr() != b.data_ptr()
    a.zero_()  # a caller's copy: the cached response is untouched
    assert torch.equal(F.ramp_response(73, "hann", 0.8), b)
    assert not torch.equal(F.ramp_response(73, "ram-lak", 0.8), b)
    with pytest.raises(RuntimeError):
        F.ramp_response(73, "nope")


def test_cosine_weights_closed_form():
    ang = np.linspace(0, 2 * np.pi, 5, endpoint=False)
    g = FanGeometry(DSD=[24.0, 25.0, 26.0, 27.0, 28.0], DSO=16.0, angles=ang, vol_shape=(8, 8),
                    det_shape=37, det_spacing=0.75)
    u = (np.arange(37) - 18.0) * 0.75
    dsd = np.array([24.0, 25.0, 26.0, 27.0, 28.0])[:, None]
    np.testing.assert_array_equal(F.cosine_weights_fan(g, "cpu").numpy(),
                                  (dsd / np.sqrt(dsd ** 2 + u[None, :] ** 2)).astype(np.float32))
    c = ConeGeometry(DSD=30.0, DSO=20.0, angles=ang[:3], vol_shape=(8, 8, 8), det_shape=(7, 11),
                     det_spacing=(1.5, 0.75))
    uu = (np.arange(11) - 5.0) * 0.75
    vv = (np.arange(7) - 3.0) * 1.5
    r2 = vv[:, None] ** 2 + uu[None, :] ** 2
    np.testing.assert_array_equal(F.cosine_weights_cone(c, "cpu").numpy(),
                                  (30.0 / np.sqrt(30.0 ** 2 + r2)).astype(np.float32))
