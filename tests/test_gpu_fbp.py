"""GPU parity of the FBP / FDK weightings and the filtered-backprojection pipeline.

Tolerances:
  fan backprojection, FBP weight (w^2)       rel-L2 <= 1e-5, max-abs <= 2e-5 * max|ref|
  cone backprojection, FDK weight            rel-L2 <= 1e-5, max-abs <= 2e-5 * max|ref|
  fbp() FP32 end to end vs FP64 pipeline     rel-L2 <= 1e-4 (FP32 FFT filtering: the ramp
                                             amplifies the filter's rounding)
  fbp() FP16 storage vs FP64 pipeline        rel-L2 <= 1e-3 (FP32 projections in, filtered
                                             projections and volume stored in FP16)
  fbp() BF16 storage vs FP64 pipeline        rel-L2 <= 5e-3 (8-bit significand: rounding the
                                             stored volume alone costs ~1.5e-3)
Rounding the raw projections to FP16 before the ramp filter costs ~1.5e-3 by itself (the
filter amplifies the rounding noise), so the 16-bit paths filter in FP32 and store the filtered
projections. The FP64 pipeline is curadon_amd.reconstruction.fbp_filter on FP64 host tensors followed by the oracle's
weighted backprojector.
"""
import numpy as np
import pytest
import torch

import curadon_amd as cm
from curadon_amd import reconstruction as F
from curadon_amd import phantom
from geoms import as_cone_geometry, as_fan_geometry, c2, cone, fan

pytestmark = pytest.mark.gpu


def rel_l2(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def check_close(got, ref, what, rtol=1e-5, maxabs=2e-5):
    r = rel_l2(got, ref)
    m = float(np.abs(got - ref).max() / max(np.abs(ref).max(), 1e-300))
    print(f"{what}: rel-L2 {r:.3e}  max-abs/max {m:.3e}", flush=True)
    assert r <= rtol, (what, r)
    assert m <= maxabs, (what, m)


def dev(x, gpu, dtype=torch.float32):
    return torch.as_tensor(np.ascontiguousarray(x)).to(device=gpu, dtype=dtype)


def host(t):
    return t.double().cpu().numpy()


def test_fan_fbp_weighting_small(oracle, gpu):
    rec = fan(64, 91, 90, 128.0, 192.0)
    geom = as_fan_geometry(rec)
    y = np.random.default_rng(20).random((90, 91))
    check_close(host(cm.backward(dev(y, gpu), geom, mode="fbp")), oracle.bp2d(rec, y, wpow=2),
                "fan bp2d fbp weighting 64^2")


def test_fan_fbp_weighting_c2_rows(oracle, gpu):
    rec = c2()
    geom = as_fan_geometry(rec)
    y = np.random.default_rng(21).random((720, 1449))
    out = host(cm.backward(dev(y, gpu), geom, mode="fbp"))
    rows = (0, 300, 511, 1023)
    ref = np.stack([oracle.bp2d(rec, y, rows=(r, r + 1), wpow=2)[r] for r in rows])
    check_close(out[list(rows)], ref, "c2 bp2d fbp weighting (4 rows)")


def test_cone_fdk_weighting(oracle, gpu):
    rec = cone(32, 48, 40, 24, 64.0, 96.0)
    geom = as_cone_geometry(rec)
    y = np.random.default_rng(22).random((24, 40, 48))
    check_close(host(cm.backward(dev(y, gpu), geom, mode="fdk")),
                oracle.bp3d(rec, y, fdk=True), "cone bp3d fdk weighting")


def _fan_problem(n=256, A=512):
    rec = fan(n, int(np.ceil(np.sqrt(2) * n)), A, 2.0 * n, 3.0 * n)
    img = phantom.ellipses_2d(n, seed=0)
    return rec, img


def test_fbp_pipeline_fp32(oracle, gpu):
    rec, img = _fan_problem()
    geom = as_fan_geometry(rec)
    sino = oracle.fwd2d(rec, img)
    q64 = F.fbp_filter(torch.tensor(sino), geom).numpy()
    ref = oracle.bp2d(rec, q64, wpow=2)
    got = host(cm.fbp(dev(sino, gpu), geom))
    r = rel_l2(got, ref)
    print(f"fbp FP32 pipeline 256^2: rel-L2 {r:.3e}")
    assert r <= 1e-4


@pytest.mark.parametrize("prec,tol", [(torch.float16, 1e-3), (torch.bfloat16, 5e-3)])
def test_fbp_pipeline_half_storage(oracle, gpu, prec, tol):
    rec, img = _fan_problem()
    geom = as_fan_geometry(rec, vol_prec=prec, sino_prec=prec)
    sino = oracle.fwd2d(rec, img)
    q64 = F.fbp_filter(torch.tensor(sino), geom).numpy()
    ref = oracle.bp2d(rec, q64, wpow=2)
    out = cm.fbp(dev(sino, gpu), geom)
    assert out.dtype == prec
    r = rel_l2(host(out), ref)
    print(f"fbp {prec} storage 256^2: rel-L2 {r:.3e}")
    assert r <= tol


def test_fdk_pipeline_fp32(oracle, gpu):
    rec = cone(48, 72, 72, 96, 96.0, 144.0)
    geom = as_cone_geometry(rec)
    vol = phantom.ellipsoids_3d(48, seed=2, dtype=np.float64)
    sino = oracle.fwd3d(rec, vol)
    q64 = F.fbp_filter(torch.tensor(sino), geom).numpy()
    ref = oracle.bp3d(rec, q64, fdk=True)
    got = host(cm.fbp(dev(sino, gpu), geom))
    r = rel_l2(got, ref)
    print(f"fdk FP32 pipeline 48^3: rel-L2 {r:.3e}")
    assert r <= 1e-4


@pytest.mark.parametrize("shape,cw_rows,window", [((7, 37), 7, "ram-lak"), ((3, 5, 64), 5, "hann"),
                                                  ((2, 9, 33), 9, "shepp-logan"),
                                                  ((5, 1100), 5, "hamming"),
                                                  ((3, 1025), 1, "cosine"),
                                                  ((3, 2100), 3, "ram-lak")])
def test_pair_filter_matches_real_transforms(gpu, shape, cw_rows, window):
    """The device filter against torch.fft's real transforms in FP64 on the same weighted rows:
    the Toeplitz product (rows up to 1024 cells), the native pair filter (two rows per complex
    FFT; L = 2250, 2160) with odd and even row counts, batches and windows, and rows long enough
    (L = 4320) to take the real-transform branch."""
    g = torch.Generator().manual_seed(3)
    p = torch.rand(shape, generator=g, dtype=torch.float64) - 0.3
    D = shape[-1]
    cw = torch.rand((cw_rows, D), generator=g, dtype=torch.float64) + 0.5
    rep = p.numel() // D // cw_rows
    weighted = (p.reshape(rep, cw_rows, D) * cw).reshape(shape)
    ref = F.filter_rows(weighted, F.ramp_response(D, window, 0.9).double())
    got = F.filter_rows_weighted(p.float().to(gpu), cw.float().to(gpu), D, window, 0.9)
    assert got.shape == p.shape and got.dtype == torch.float32
    err = (got.double().cpu() - ref).norm() / ref.norm()
    assert err < 1e-6, err


def test_toeplitz_filter_ignores_tf32_setting(gpu):
    """The Toeplitz product stays in full FP32 when the caller allows TF32 matmuls: the same
    bits as with TF32 off, and the caller's setting is restored."""
    g = torch.Generator().manual_seed(5)
    p = (torch.rand((6, 300), generator=g) - 0.5).to(gpu)
    cw = (torch.rand((6, 300), generator=g) + 0.5).to(gpu)
    ref = F.filter_rows_weighted(p, cw, 300)
    old = torch.backends.cuda.matmul.allow_tf32
    torch.backends.cuda.matmul.allow_tf32 = True
    try:
        got = F.filter_rows_weighted(p, cw, 300)
        assert torch.backends.cuda.matmul.allow_tf32 is True
    finally:
        torch.backends.cuda.matmul.allow_tf32 = old
    assert torch.equal(got, ref)
