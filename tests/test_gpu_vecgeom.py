"""GPU parity of plans built from ASTRA-like views (FanGeometry.from_vec ->
cmi_plan2d_create_vec) against the FP64 oracle's vector form: a moved source, a shifted and
tilted detector and a per-view cell spacing (no counterpart in the reference, whose from_vec
keeps only DSO / DSD / angle, python/standalone/src/curadon/geometry.py:84-133).

Tolerances as for circular plans (FP32 vs FP64): rel-L2 <= 1e-5, max-abs <= 2e-5 max|ref|;
dot test 1e-5. Views from a circular geometry reproduce its operators to FP32 rounding."""
from types import SimpleNamespace

import numpy as np
import pytest
import torch

import curadon_amd as cm
from curadon_amd.geometry import FanGeometry
from geoms import fan

pytestmark = pytest.mark.gpu


def rel_l2(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def check(got, ref, what, rtol=1e-5, maxabs=2e-5):
    r = rel_l2(got, ref)
    m = float(np.abs(got - ref).max() / np.abs(ref).max())
    print(f"{what}: rel-L2 {r:.3e} max-abs/max {m:.3e}", flush=True)
    assert r <= rtol and m <= maxabs, (what, r, m)


def general_views(seed, A, dso, dsd):
    r = np.random.default_rng(seed)
    out = []
    for th in r.uniform(0, 2 * np.pi, A):
        c, s = np.cos(th), np.sin(th)
        R = np.array([[c, s], [-s, c]])
        tilt, du = r.uniform(-0.35, 0.35), r.uniform(0.7, 1.4)
        out.append(np.concatenate([R @ [r.uniform(-3, 3), dso], R @ [r.uniform(-4, 4), dso - dsd],
                                   R @ (du * np.array([np.cos(tilt), np.sin(tilt)]))]))
    return np.array(out)


def record(geom):
    """The oracle's view of a from_vec geometry."""
    return SimpleNamespace(vol_shape=geom.vol_shape, vol_spacing=geom.vol_spacing,
                           vol_offset=geom.vol_offset, det_shape=geom.det_shape,
                           det_spacing=geom.det_spacing, det_offset=0.0, angles=geom.angles,
                           DSO=geom.DSO, DSD=geom.DSD, vecs=geom.vecs)


@pytest.fixture(scope="module")
def vgeom():
    views = general_views(3, 40, 120.0, 200.0)
    return FanGeometry.from_vec((48, 64), 101, views, vol_spacing=(1.1, 0.9),
                                vol_offset=(0.7, -0.4))


def test_views_forward_backward_adjoint(oracle, vgeom, gpu):
    rec = record(vgeom)
    x = np.random.default_rng(4).random((48, 64))
    check(cm.forward(torch.tensor(x, dtype=torch.float32, device=gpu), vgeom).double().cpu().numpy(),
          oracle.fwd2d(rec, x), "views fwd2d")
    y = np.random.default_rng(5).random((40, 101))
    yt = torch.tensor(y, dtype=torch.float32, device=gpu)
    check(cm.backward(yt, vgeom).double().cpu().numpy(), oracle.bp2d(rec, y), "views bp2d")
    check(cm.backward(yt, vgeom, mode="fbp").double().cpu().numpy(), oracle.bp2d(rec, y, wpow=2),
          "views bp2d fbp weight")
    check(cm.adjoint(yt, vgeom).double().cpu().numpy(), oracle.adj2d(rec, y), "views adj2d")
    xt = torch.tensor(x, dtype=torch.float32, device=gpu)
    lhs = float(torch.vdot(cm.forward(xt, vgeom).double().ravel(), yt.double().ravel()))
    rhs = float(torch.vdot(xt.double().ravel(), cm.adjoint(yt, vgeom).double().ravel()))
    print(f"views dot test {abs(lhs - rhs) / abs(lhs):.3e}")
    assert abs(lhs - rhs) / abs(lhs) <= 1e-5


def test_views_of_a_circular_geometry_reproduce_it(gpu):
    g = cm.FanGeometry(DSD=200.0, DSO=130.0, angles=np.linspace(0, 2 * np.pi, 60, endpoint=False),
                       vol_shape=(64, 64), det_shape=91, det_spacing=1.3, det_offset=0.0)
    v = FanGeometry.from_vec((64, 64), 91, g.to_vec())
    x = torch.rand(64, 64, device=gpu)
    y = torch.rand(60, 91, device=gpu)
    for name, a, b in (("fwd", cm.forward(x, v), cm.forward(x, g)),
                       ("bp", cm.backward(y, v), cm.backward(y, g)),
                       ("adj", cm.adjoint(y, v), cm.adjoint(y, g))):
        r = rel_l2(a.double().cpu().numpy(), b.double().cpu().numpy())
        print(f"circular views {name}: rel-L2 {r:.3e}")
        assert r <= 2e-6, name


def test_views_refused_by_fbp_and_serving(vgeom, gpu):
    from curadon_amd.serving import FBPJob, FBPServer
    y = torch.rand(40, 101, device=gpu)
    with pytest.raises(NotImplementedError):
        cm.fbp(y, vgeom)
    with pytest.raises(NotImplementedError):
        FBPServer().run([FBPJob(geom=vgeom, projections=y)])


# ------------------------------------------- exact detector offset, rotation and COR -------

def _exact_pair(det_offset, rot, cor, n=48, D=71, A=40, dso=96.0, dsd=150.0, du=1.1):
    rec = fan(n, D, A, dso, dsd, du=du, det_offset=det_offset, spacing=(1.0, 1.2),
              offset=(0.6, -0.4), ny=44)
    rec.exact_geometry, rec.COR, rec.det_rotation = True, cor, rot
    geom = FanGeometry(DSD=rec.DSD, DSO=rec.DSO, angles=rec.angles, vol_shape=rec.vol_shape,
                       det_shape=D, det_spacing=du, det_offset=det_offset, det_rotation=rot,
                       vol_spacing=rec.vol_spacing, vol_offset=rec.vol_offset, COR=cor,
                       exact_geometry=True)
    return rec, geom


@pytest.mark.parametrize("det_offset,rot,cor", [(3.3, 0.0, 0.0), (0.0, 0.06, 0.0),
                                                (0.0, 0.0, 1.7), (-2.2, -0.05, 2.4)])
def test_exact_geometry_against_oracle(oracle, gpu, det_offset, rot, cor):
    """Circular plans with exact detector geometry (cmi_plan2d_create_ex): the forward, the
    voxel-driven backprojection and the matched adjoint follow the shifted, turned detector
    and the moved rotation centre, against the oracle's views of the same geometry; the
    backprojection's footprint moves with det_offset like the forward's (dot test of the
    forward against the adjoint, and the backprojection against the oracle's, which shifts the
    detector coordinate)."""
    rec, geom = _exact_pair(det_offset, rot, cor)
    img = np.random.default_rng(11).random((44, 48))
    x = torch.tensor(img, dtype=torch.float32, device=gpu)
    check(cm.forward(x, geom).double().cpu().numpy(), oracle.fwd2d(rec, img),
          f"exact fwd (off {det_offset}, rot {rot}, cor {cor})")
    y = np.random.default_rng(12).random((40, 71))
    yt = torch.tensor(y, dtype=torch.float32, device=gpu)
    check(cm.backward(yt, geom).double().cpu().numpy(), oracle.bp2d(rec, y),
          f"exact bp (off {det_offset}, rot {rot}, cor {cor})")
    at = cm.adjoint(yt, geom)
    check(at.double().cpu().numpy(), oracle.adj2d(rec, y),
          f"exact adj (off {det_offset}, rot {rot}, cor {cor})")
    lhs = torch.vdot(cm.forward(x, geom).double().ravel(), yt.double().ravel())
    rhs = torch.vdot(x.double().ravel(), at.double().ravel())
    assert abs(float(lhs - rhs)) / abs(float(lhs)) <= 1e-5


def test_exact_geometry_without_offsets_is_the_reference_projector(gpu):
    """Exact geometry with no offset, rotation or COR projects like the default plan (the
    same rays; the forward to FP64 set-up rounding, the backprojection through the view form
    to FP32 rounding)."""
    rec, ex = _exact_pair(0.0, 0.0, 0.0)
    plain = FanGeometry(DSD=rec.DSD, DSO=rec.DSO, angles=rec.angles, vol_shape=rec.vol_shape,
                        det_shape=71, det_spacing=1.1, vol_spacing=rec.vol_spacing,
                        vol_offset=rec.vol_offset)
    x = torch.rand(44, 48, device=gpu, generator=torch.Generator(device=gpu).manual_seed(1))
    check(cm.forward(x, ex).double().cpu().numpy(), cm.forward(x, plain).double().cpu().numpy(),
          "exact (no offsets) vs default fwd", rtol=1e-6, maxabs=1e-5)
    y = torch.rand(40, 71, device=gpu, generator=torch.Generator(device=gpu).manual_seed(2))
    check(cm.backward(y, ex).double().cpu().numpy(), cm.backward(y, plain).double().cpu().numpy(),
          "exact (no offsets) vs default bp", rtol=1e-6, maxabs=1e-5)
