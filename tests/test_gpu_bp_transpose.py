"""GPU parity of the transpose of the fan-beam backprojection (cmi_backward_2d_transpose,
``backward_transpose``), the gradient that TorchBackprojection now uses for fan beams, against
the FP64 oracle's transpose (``oracle.bp2d_T``, pinned against the assembled matrix of
``oracle.bp2d``).

Tolerances (FP32 positions and weights vs FP64): rel-L2 <= 1e-5, max-abs <= 2e-5 max|ref|;
dot tests against the GPU backprojection 1e-5 (max-abs 5e-5 at 1024^2, see there). At 1024^2 the oracle runs on projection
subsets (each projection's row of B^T is independent of the others)."""
import numpy as np
import pytest
import torch

import curadon_amd as cm
from curadon_amd.geometry import FanGeometry
from geoms import as_fan_geometry, c1, c2, fan
from test_gpu_vecgeom import general_views, record

pytestmark = pytest.mark.gpu


def rel_l2(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def check(got, ref, what, rtol=1e-5, maxabs=2e-5):
    r = rel_l2(got, ref)
    m = float(np.abs(got - ref).max() / max(np.abs(ref).max(), 1e-300))
    print(f"{what}: rel-L2 {r:.3e}  max-abs/max {m:.3e}", flush=True)
    assert r <= rtol, (what, r)
    assert m <= maxabs, (what, m)


def dev(x, gpu, dtype=torch.float32):
    return torch.as_tensor(np.ascontiguousarray(x)).to(device=gpu, dtype=dtype)


def host(t):
    return t.double().cpu().numpy()


@pytest.mark.parametrize("mode,wpow", [("reference", 1), ("fbp", 2)])
def test_c1_transpose(oracle, gpu, mode, wpow):
    rec = c1()
    geom = as_fan_geometry(rec)
    v = np.random.default_rng(40).random((64, 64)) - 0.2
    got = host(cm.backward_transpose(dev(v, gpu), geom, mode=mode))
    check(got, oracle.bp2d_T(rec, v, wpow=wpow), f"c1 B^T ({mode})")


def test_offsets_and_spacings(oracle, gpu):
    rec = fan(48, 70, 90, 100.0, 160.0, du=1.3, det_offset=0.6, spacing=(1.1, 0.9),
              offset=(0.7, -0.4), ny=40)
    geom = as_fan_geometry(rec)
    v = np.random.default_rng(41).random((40, 48))
    check(host(cm.backward_transpose(dev(v, gpu), geom)), oracle.bp2d_T(rec, v),
          "B^T with offsets and spacings")


@pytest.mark.parametrize("mode,wpow", [("reference", 1), ("fbp", 2)])
def test_views(oracle, gpu, mode, wpow):
    geom = FanGeometry.from_vec((48, 64), 101, general_views(3, 40, 120.0, 200.0),
                                vol_spacing=(1.1, 0.9), vol_offset=(0.7, -0.4))
    v = np.random.default_rng(42).random((48, 64))
    check(host(cm.backward_transpose(dev(v, gpu), geom, mode=mode)),
          oracle.bp2d_T(record(geom), v, wpow=wpow), f"views B^T ({mode})")


def test_odd_shapes_partial_segments(oracle, gpu):
    """7 cells (one partial segment), a 9 x 6 image, batch of 2."""
    rec = fan(9, 7, 3, 30.0, 50.0, ny=6)
    geom = as_fan_geometry(rec)
    v = np.random.default_rng(47).random((2, 6, 9))
    for mode, wpow in (("reference", 1), ("fbp", 2)):
        got = host(cm.backward_transpose(dev(v, gpu), geom, mode=mode))
        for i in range(2):
            check(got[i], oracle.bp2d_T(rec, v[i], wpow=wpow), f"odd shapes B^T ({mode}) image {i}")


def test_wide_cells_and_wide_wedges(oracle, gpu):
    """Detector cells so wide that a segment's boundary lines meet the image lines at grazing
    angles (the workgroups walk whole lines), and coarse cells whose 64-cell segments cover most
    of the image."""
    for rec, what in ((fan(16, 8, 6, 40.0, 50.0, du=20.0), "20-unit cells"),
                      (fan(48, 100, 30, 60.0, 90.0, du=3.0, ny=40), "3-unit cells")):
        v = np.random.default_rng(46).random(tuple(int(x) for x in rec.vol_shape))
        for mode, wpow in (("reference", 1), ("fbp", 2)):
            got = host(cm.backward_transpose(dev(v, gpu), as_fan_geometry(rec), mode=mode))
            check(got, oracle.bp2d_T(rec, v, wpow=wpow), f"B^T {what} ({mode})")


@pytest.mark.parametrize("a0,a1", [(0, 4), (178, 182), (357, 361), (700, 704)])
def test_c2_projection_blocks(oracle, gpu, a0, a1):
    """The 1024^2 workload, full GPU transpose, oracle on four projections per block. Max-abs
    bar 5e-5: the FP32 detector coordinate near u = 1449 carries about 1e-4 cells of rounding,
    which moves the split of each pixel between its two cells by as much, and a cell sums about
    a thousand such pixels of one projection with correlated roundings (rel-L2 stays near
    2e-6), as for the adjoint at this size."""
    rec = c2()
    geom = as_fan_geometry(rec)
    v = np.random.default_rng(43 + a0).random((1024, 1024))
    got = host(cm.backward_transpose(dev(v, gpu), geom))[a0:a1]
    ref = oracle.bp2d_T(rec, v, angles=(a0, a1))[a0:a1]
    check(got, ref, f"c2 B^T (projections {a0}..{a1 - 1})", maxabs=5e-5)


def test_c2_dot_product_with_backprojection(gpu):
    geom = as_fan_geometry(c2())
    g = torch.Generator(device="cpu").manual_seed(44)
    s = torch.rand(720, 1449, generator=g).to(gpu)
    v = (torch.rand(1024, 1024, generator=g) - 0.3).to(gpu)
    for mode in ("reference", "fbp"):
        lhs = torch.vdot(cm.backward(s, geom, mode=mode).double().ravel(), v.double().ravel())
        rhs = torch.vdot(s.double().ravel(),
                         cm.backward_transpose(v, geom, mode=mode).double().ravel())
        r = abs(float(lhs - rhs)) / abs(float(lhs))
        print(f"c2 B / B^T dot test ({mode}): {r:.3e}")
        assert r <= 1e-5


def test_batched_equals_single_and_is_deterministic(gpu):
    geom = as_fan_geometry(c1())
    v = torch.rand(3, 64, 64, device=gpu)
    batched = cm.backward_transpose(v, geom)
    for i in range(3):
        assert torch.equal(batched[i], cm.backward_transpose(v[i], geom))
    assert torch.equal(batched, cm.backward_transpose(v, geom))


@pytest.mark.parametrize("prec,tol", [(torch.float16, 1e-3), (torch.bfloat16, 6e-3)])
def test_half_storage(oracle, gpu, prec, tol):
    rec = c1()
    geom = as_fan_geometry(rec, vol_prec=prec, sino_prec=prec)
    v = dev(np.random.default_rng(45).random((64, 64)), gpu, prec)
    got = host(cm.backward_transpose(v, geom))
    check(got, oracle.bp2d_T(rec, host(v)), f"c1 B^T {prec}", rtol=tol, maxabs=4 * tol)


def test_torch_backprojection_gradient_is_the_transpose(gpu):
    geom = as_fan_geometry(c1())
    s = torch.rand(360, 91, device=gpu, requires_grad=True)
    v = torch.rand(64, 64, device=gpu)
    out = cm.TorchBackprojection.apply(s, geom)
    (out * v).sum().backward()
    assert torch.equal(s.grad, cm.backward_transpose(v, geom))


def test_zero_image_gives_zero_sinogram(gpu):
    geom = as_fan_geometry(c1())
    out = cm.backward_transpose(torch.zeros(64, 64, device=gpu), geom)
    assert not out.any()


def test_source_inside_the_image_is_refused(gpu):
    """A source inside the image's bounding box leaves the weight unbounded."""
    geom = cm.FanGeometry(DSD=60.0, DSO=20.0, angles=np.linspace(0, 2 * np.pi, 8, endpoint=False),
                          vol_shape=(64, 64), det_shape=91)
    with pytest.raises(NotImplementedError):
        cm.backward_transpose(torch.rand(64, 64, device=gpu), geom)
