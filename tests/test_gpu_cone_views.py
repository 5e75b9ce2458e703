"""GPU parity of cone-beam plans from ASTRA-like views (cmi_plan3d_create_vec,
ConeGeometry.from_vec) against the FP64 oracle's vector path, through the C-ABI: a helical,
non-circular trajectory with a tilted panel of non-square pixels, anisotropic voxels and a
volume offset. Tolerances as for the Euler-angle plans: rel-L2 <= 1e-5, max-abs <= 2e-5 max|ref|;
dot test 1e-5."""
import numpy as np
import pytest
import torch

import curadon_amd as cm
from curadon_amd import phantom
from geoms import as_cone_geometry, cone, helix_views as _helix, with_views as _with_views

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


@pytest.fixture(scope="module")
def helix(gpu):
    rec = cone(28, 40, 34, 18, 60.0, 100.0, vol_spacing=(1.1, 0.9, 1.0),
               vol_offset=(0.5, -0.7, 0.3), nz=24, ny=26)
    vecs = _helix(18, 28, 60.0, 20.0, 40.0, 40, 34, du=1.2, dv=0.9, tilt=0.12)
    rv = _with_views(rec, vecs)
    geom = cm.ConeGeometry.from_vec(rec.vol_shape, rec.det_shape, vecs,
                                    vol_spacing=rec.vol_spacing, vol_offset=rec.vol_offset)
    return rv, geom


def test_views_forward(oracle, helix, gpu):
    rv, geom = helix
    vol = phantom.ellipsoids_3d(28, seed=4, dtype=np.float64)[:24, :26, :]
    got = host(cm.forward(dev(vol, gpu), geom))
    check_close(got, oracle.fwd3d(rv, vol), "helix views fwd3d")
    assert geom.plan.info()["total_samples"] == oracle.fwd3d(rv, vol, count=True)[1]


def test_views_backprojection_and_adjoint(oracle, helix, gpu):
    rv, geom = helix
    y = np.random.default_rng(3).random((18, 34, 40))
    yt = dev(y, gpu)
    ref_m = oracle.bp3d(rv, y, matched=True)
    check_close(host(cm.backward(yt, geom, mode="matched")), ref_m, "helix views bp3d matched")
    # reference mode of a plan from views projects like matched mode
    assert torch.equal(cm.backward(yt, geom), cm.backward(yt, geom, mode="matched"))
    with pytest.raises(NotImplementedError):
        cm.backward(yt, geom, mode="fdk")
    check_close(host(cm.adjoint(yt, geom)), oracle.adj3d(rv, y), "helix views adj3d")


def test_views_dot_product_and_batch(helix, gpu):
    rv, geom = helix
    g = torch.Generator(device=gpu).manual_seed(7)
    x = torch.rand(2, 24, 26, 28, device=gpu, generator=g)
    y = torch.rand(2, 18, 34, 40, device=gpu, generator=g)
    ax = cm.forward(x, geom)
    aty = cm.adjoint(y, geom)
    lhs = torch.vdot(ax.double().ravel(), y.double().ravel())
    rhs = torch.vdot(x.double().ravel(), aty.double().ravel())
    r = abs(float(lhs - rhs)) / abs(float(lhs))
    print(f"helix views dot test: {r:.3e}")
    assert r <= 1e-5
    for i in range(2):
        assert torch.equal(ax[i], cm.forward(x[i].contiguous(), geom))
        assert torch.equal(aty[i], cm.adjoint(y[i].contiguous(), geom))


def test_views_of_euler_geometry_project_alike(gpu):
    """from_vec(to_vec(g)) of a tilted-orbit Euler geometry with a rotated panel and COR
    projects like g (the frames agree to FP64 rounding, so only rays whose step count sits on a
    rounding edge could differ)."""
    r = np.random.default_rng(5)
    ang = np.vstack([r.uniform(0, 2 * np.pi, 8), r.uniform(-0.3, 0.3, 8), r.uniform(-0.3, 0.3, 8)])
    rec = cone(20, 30, 26, 8, 50.0, 80.0, du=(1.2, 0.9), angles=ang, vol_spacing=(1.1, 0.9, 1.0),
               vol_offset=(0.5, -0.7, 0.3), det_rotation=(0.04, -0.06, 0.03), cor=0.8, nz=18,
               ny=22)
    g = as_cone_geometry(rec)
    gv = cm.ConeGeometry.from_vec(g.vol_shape, g.det_shape, g.to_vec(),
                                  vol_spacing=g.vol_spacing, vol_offset=g.vol_offset)
    x = torch.rand(18, 22, 20, device=gpu, generator=torch.Generator(device=gpu).manual_seed(1))
    a, b = host(cm.forward(x, g)), host(cm.forward(x, gv))
    check_close(b, a, "from_vec(to_vec) fwd3d", rtol=1e-6, maxabs=1e-5)
    y = torch.rand(8, 26, 30, device=gpu, generator=torch.Generator(device=gpu).manual_seed(2))
    check_close(host(cm.backward(y, gv, mode="matched")), host(cm.backward(y, g, mode="matched")),
                "from_vec(to_vec) bp3d matched", rtol=1e-6, maxabs=1e-5)


def test_views_refused_where_unsupported(helix, gpu):
    _, geom = helix
    with pytest.raises(NotImplementedError):
        cm.fbp(torch.zeros(18, 34, 40, device=gpu), geom)
