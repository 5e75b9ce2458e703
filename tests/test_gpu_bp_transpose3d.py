"""GPU parity of the transpose of the cone-beam backprojection (cmi_backward_3d_transpose,
``backward_transpose``), the gradient that TorchBackprojection uses for cone beams, against the
FP64 oracle's transpose (``oracle.bp3d_T``, pinned against the assembled matrix of
``oracle.bp3d``).

Tolerances (FP32 matrices and weights vs FP64): rel-L2 <= 1e-5, max-abs <= 2e-5 max|ref|;
dot tests against the GPU backprojection 1e-5. At 256^3 the oracle runs on projection subsets
(each projection of B^T is independent of the others)."""
import numpy as np
import pytest
import torch

import curadon_amd as cm
from geoms import as_cone_geometry, c3, cone, helix_views, with_views

pytestmark = pytest.mark.gpu

MODES = [("reference", {}), ("matched", {"matched": True}),
         ("fdk", {"matched": True, "fdk": True})]


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


def small():
    return cone(32, 48, 40, 24, 64.0, 96.0, du=(1.3, 1.1))


@pytest.mark.parametrize("mode,kw", MODES)
def test_small_all_modes(oracle, gpu, mode, kw):
    rec = small()
    v = np.random.default_rng(50).random((32, 32, 32)) - 0.2
    got = host(cm.backward_transpose(dev(v, gpu), as_cone_geometry(rec), mode=mode))
    check(got, oracle.bp3d_T(rec, v, **kw), f"32^3 B^T ({mode})")


@pytest.mark.parametrize("mode,kw", MODES[:2])
def test_general_geometry(oracle, gpu, mode, kw):
    """Euler angles off the circle, a turned and offset panel, anisotropic voxels, COR shift:
    projection matrices with every entry non-zero, march axes x, y and z."""
    r = np.random.default_rng(1)
    ang = np.vstack([r.uniform(0, 2 * np.pi, 10), r.uniform(-0.3, 0.3, 10),
                     r.uniform(-0.3, 0.3, 10)])
    rec = cone(24, 36, 30, 10, 60.0, 100.0, du=(1.2, 0.9), angles=ang,
               vol_spacing=(1.1, 0.9, 1.0), vol_offset=(0.5, -0.7, 0.3), det_offset=(1.5, -2.0),
               det_rotation=(0.2, -0.1, 0.15), cor=0.8, nz=20, ny=22)
    v = np.random.default_rng(51).random((20, 22, 24))
    got = host(cm.backward_transpose(dev(v, gpu), as_cone_geometry(rec), mode=mode))
    check(got, oracle.bp3d_T(rec, v, **kw), f"general geometry B^T ({mode})")


def test_wide_pixels_walk_whole_slices(oracle, gpu):
    """Pixels so wide that a tile's corner rays meet the slices at grazing angles: the
    workgroups walk whole slices instead of the corner lines' boxes."""
    rec = cone(16, 8, 8, 6, 40.0, 50.0, du=(20.0, 20.0))
    v = np.random.default_rng(52).random((16, 16, 16))
    for mode, kw in MODES:
        got = host(cm.backward_transpose(dev(v, gpu), as_cone_geometry(rec), mode=mode))
        check(got, oracle.bp3d_T(rec, v, **kw), f"wide pixels B^T ({mode})")


def test_odd_shapes_partial_tiles(oracle, gpu):
    """A 7 x 5 panel (one partial tile) and a 9 x 8 x 7 volume, batch of 2."""
    rec = cone(9, 7, 5, 3, 30.0, 50.0, nz=7, ny=8)
    geom = as_cone_geometry(rec)
    v = np.random.default_rng(57).random((2, 7, 8, 9))
    for mode, kw in MODES:
        got = host(cm.backward_transpose(dev(v, gpu), geom, mode=mode))
        for i in range(2):
            check(got[i], oracle.bp3d_T(rec, v[i], **kw), f"odd shapes B^T ({mode}) volume {i}")


def test_views(oracle, gpu):
    rec = cone(28, 40, 34, 18, 60.0, 100.0, vol_spacing=(1.1, 0.9, 1.0),
               vol_offset=(0.5, -0.7, 0.3), nz=24, ny=26)
    vecs = helix_views(18, 28, 60.0, 20.0, 40.0, 40, 34, du=1.2, dv=0.9, tilt=0.12)
    rv = with_views(rec, vecs)
    geom = cm.ConeGeometry.from_vec(rec.vol_shape, rec.det_shape, vecs,
                                    vol_spacing=rec.vol_spacing, vol_offset=rec.vol_offset)
    v = dev(np.random.default_rng(53).random((24, 26, 28)), gpu)
    got = cm.backward_transpose(v, geom, mode="matched")
    check(host(got), oracle.bp3d_T(rv, host(v), matched=True), "helix views B^T")
    assert torch.equal(cm.backward_transpose(v, geom), got)  # reference mode = matched
    with pytest.raises(NotImplementedError):
        cm.backward_transpose(v, geom, mode="fdk")


@pytest.mark.parametrize("a0,a1", [(0, 3), (44, 47), (90, 93), (301, 304)])
def test_c3_projection_blocks(oracle, gpu, a0, a1):
    """The 256^3 workload (384^2 panel, 360 projections), full GPU transpose, oracle on three
    projections per block (march axes x and z, 45 degrees between them)."""
    rec = c3()
    v = np.random.default_rng(54 + a0).random((256, 256, 256)).astype(np.float32)
    vt = dev(v, gpu)
    for mode, kw in (MODES[0], MODES[2]):
        got = host(cm.backward_transpose(vt, as_cone_geometry(rec), mode=mode))[a0:a1]
        ref = oracle.bp3d_T(rec, v.astype(np.float64), angles=(a0, a1), **kw)[a0:a1]
        check(got, ref, f"c3 B^T {mode} (projections {a0}..{a1 - 1})")


def test_c3_dot_product_with_backprojection(gpu):
    geom = as_cone_geometry(c3())
    g = torch.Generator(device="cpu").manual_seed(55)
    s = torch.rand(360, 384, 384, generator=g).to(gpu)
    v = (torch.rand(256, 256, 256, generator=g) - 0.3).to(gpu)
    for mode in ("reference", "matched", "fdk"):
        lhs = torch.vdot(cm.backward(s, geom, mode=mode).double().ravel(), v.double().ravel())
        rhs = torch.vdot(s.double().ravel(),
                         cm.backward_transpose(v, geom, mode=mode).double().ravel())
        r = abs(float(lhs - rhs)) / abs(float(lhs))
        print(f"c3 B / B^T dot test ({mode}): {r:.3e}", flush=True)
        assert r <= 1e-5


def test_batched_equals_single_and_is_deterministic(gpu):
    geom = as_cone_geometry(small())
    v = torch.rand(3, 32, 32, 32, device=gpu)
    batched = cm.backward_transpose(v, geom, mode="fdk")
    for i in range(3):
        assert torch.equal(batched[i], cm.backward_transpose(v[i], geom, mode="fdk"))
    assert torch.equal(batched, cm.backward_transpose(v, geom, mode="fdk"))


@pytest.mark.parametrize("prec,tol", [(torch.float16, 1e-3), (torch.bfloat16, 6e-3)])
def test_half_storage(oracle, gpu, prec, tol):
    rec = small()
    geom = as_cone_geometry(rec, vol_prec=prec, sino_prec=prec)
    v = dev(np.random.default_rng(56).random((32, 32, 32)), gpu, prec)
    got = host(cm.backward_transpose(v, geom))
    check(got, oracle.bp3d_T(rec, host(v)), f"32^3 B^T {prec}", rtol=tol, maxabs=4 * tol)


def test_torch_backprojection_gradient_is_the_transpose(gpu):
    geom = as_cone_geometry(small())
    s = torch.rand(24, 40, 48, device=gpu, requires_grad=True)
    v = torch.rand(32, 32, 32, device=gpu)
    out = cm.TorchBackprojection.apply(s, geom)
    (out * v).sum().backward()
    assert torch.equal(s.grad, cm.backward_transpose(v, geom))


def test_zero_volume_gives_zero_sinogram(gpu):
    geom = as_cone_geometry(small())
    assert not cm.backward_transpose(torch.zeros(32, 32, 32, device=gpu), geom).any()


def test_source_inside_the_volume_is_refused(gpu):
    """A source inside the volume leaves W <= 0 at some voxel; TorchBackprojection's gradient
    then falls back to the forward projection."""
    rec = cone(32, 48, 40, 8, 10.0, 40.0)
    geom = as_cone_geometry(rec)
    v = torch.rand(32, 32, 32, device=gpu)
    with pytest.raises(NotImplementedError):
        cm.backward_transpose(v, geom)
    s = torch.rand(8, 40, 48, device=gpu, requires_grad=True)
    (cm.TorchBackprojection.apply(s, geom) * v).sum().backward()
    assert torch.equal(s.grad, cm.forward(v, geom))
