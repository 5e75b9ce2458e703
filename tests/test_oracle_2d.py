"""Pins the FP64 oracle's 2D routines with checks that do not restate its formulas:
closed forms of axis-aligned rays, a library interpolation routine, the reference's own
ASTRA-like vector form of the geometry (geometry.py:54-79), analytic ellipse line integrals,
an explicitly assembled system matrix for the adjoint, and golden fixtures.
"""
from types import SimpleNamespace

import numpy as np
import pytest
from scipy import ndimage

from geoms import c1, fan, tiny
from curadon_amd.geometry import FanGeometry
from curadon_amd import phantom


def rng(seed):
    return np.random.default_rng(seed)


# ------------------------------------------------------------------ interpolation (a6) ---

def test_lerp1_known_answers(oracle):
    r = np.array([1.0, 2.0, 4.0, 8.0])
    for i in range(4):
        assert oracle.lerp1(r, i + 0.5) == r[i]                # texel centre
    assert oracle.lerp1(r, 0.0) == 0.5 * r[0]                  # half a texel outside
    assert oracle.lerp1(r, 4.0) == 0.5 * r[3]
    assert oracle.lerp1(r, 1.0) == 0.5 * (r[0] + r[1])
    for u in (-0.5, -3.0, 4.5, 9.0):                           # |u - .5 - i| >= 1: border 0
        assert oracle.lerp1(r, u) == 0.0


def test_bilerp_trilerp_match_scipy_grid_constant(oracle):
    """Linear filter with border addressing = map_coordinates(order=1, 'grid-constant') at
    texel coordinate - 0.5."""
    g = rng(0)
    img = g.random((7, 9))
    pts = g.uniform(-2.5, 11.5, size=(500, 2))
    ours = np.array([oracle.bilerp(img, x, y) for x, y in pts])
    ref = ndimage.map_coordinates(img, [pts[:, 1] - 0.5, pts[:, 0] - 0.5], order=1,
                                  mode="grid-constant", cval=0.0)
    np.testing.assert_allclose(ours, ref, rtol=0, atol=1e-13)
    vol = g.random((4, 5, 6))
    p3 = g.uniform(-2.0, 8.0, size=(300, 3))
    ours3 = np.array([oracle.trilerp(vol, x, y, z) for x, y, z in p3])
    ref3 = ndimage.map_coordinates(vol, [p3[:, 2] - 0.5, p3[:, 1] - 0.5, p3[:, 0] - 0.5],
                                   order=1, mode="grid-constant", cval=0.0)
    np.testing.assert_allclose(ours3, ref3, rtol=0, atol=1e-13)


# ------------------------------------------------------------------ plan (a1, a2) --------

@pytest.mark.parametrize("spacing", [(1.0, 1.0), (0.5, 2.0)])
def test_plan_matches_vector_form(oracle, spacing):
    """Source, detector centre and cell step agree with the reference's to_vec form
    (source R(0, DSO), detector centre R(0, -DOD), axis R(1, 0), R = [[c, s], [-s, c]]; this
    package's to_vec scales the axis by det_spacing, so that it is the cell step)."""
    g = fan(10, 13, 17, 30.0, 45.0, du=1.5, spacing=spacing, offset=(0.7, -1.3))
    g.angles = rng(1).uniform(0, 2 * np.pi, 17)
    fg = FanGeometry(DSD=g.DSD, DSO=g.DSO, angles=g.angles, vol_shape=g.vol_shape,
                     det_shape=g.det_shape, det_spacing=g.det_spacing,
                     vol_spacing=g.vol_spacing, vol_offset=g.vol_offset)
    vec = np.array(fg.to_vec())
    f = oracle.Fan(g)
    src, org, dl = f.plan()
    s = np.array([g.vol_spacing[1], g.vol_spacing[0]])  # (x, y)
    o = np.array([g.vol_offset[1], g.vol_offset[0]])
    ext = np.array([g.vol_shape[1], g.vol_shape[0]]) * s
    src_w = src * s + o - ext / 2
    np.testing.assert_allclose(src_w, vec[:, 0:2], atol=1e-12)
    centre = src_w + (org + (g.det_shape - 1) / 2 * dl) * s
    np.testing.assert_allclose(centre, vec[:, 2:4], atol=1e-11)
    np.testing.assert_allclose(dl * s, vec[:, 4:6], atol=1e-12)
    np.testing.assert_allclose(np.linalg.norm(vec[:, 4:6], axis=1), g.det_spacing, rtol=1e-14)
    # invariants: |src| = DSO, |centre - src| = DSD, |step| = du
    np.testing.assert_allclose(np.linalg.norm(src_w, axis=1), g.DSO, rtol=1e-14)
    np.testing.assert_allclose(np.linalg.norm(centre - src_w, axis=1), g.DSD, rtol=1e-14)


def test_plan_det_offset_moves_source(oracle):
    """det_offset shifts the source along the detector axis (plan_2d.hpp:76-80)."""
    g0 = fan(8, 7, 1, 16.0, 24.0, angles=[0.0])
    g1 = fan(8, 7, 1, 16.0, 24.0, angles=[0.0], det_offset=0.75)
    s0, o0, d0 = oracle.Fan(g0).plan()
    s1, o1, d1 = oracle.Fan(g1).plan()
    np.testing.assert_allclose(s1 - s0, [[0.75, 0.0]], atol=1e-14)
    np.testing.assert_allclose(s1 + o1, s0 + o0, atol=1e-14)  # cell 0 stays put
    np.testing.assert_allclose(d1, d0, atol=1e-15)


# ------------------------------------------------------------------ march (a3, a4) -------

def _axis_geom(n, theta):
    return fan(n, 7, 1, 40.0, 60.0, angles=[theta])


def test_vertical_ray_downward_closed_form(oracle):
    """theta = 0, centre cell of an odd detector: the ray runs down x = n/2 from the face
    y = n to y = -1 (N = n + 1, v = (0, -1)); every row gets weight 1 and the two columns
    beside x = n/2 get 1/2 each."""
    n = 8
    img = rng(2).random((n, n))
    g = _axis_geom(n, 0.0)
    N, e, v, scale = oracle.Fan(g).ray(0, 3)
    assert N == n + 1
    np.testing.assert_allclose(e, [n / 2, n], atol=1e-12)
    np.testing.assert_allclose(v, [0, -1], atol=0)
    assert scale == 1.0
    sino = oracle.fwd2d(g, img)
    expect = 0.5 * (img[:, n // 2 - 1] + img[:, n // 2]).sum()
    assert sino[0, 3] == pytest.approx(expect, rel=1e-14)


def test_vertical_ray_upward_closed_form(oracle):
    """theta = pi: the ray enters at y = -1 and takes N = n + 1 samples at y = -1 .. n - 1, so
    the clip box [-1, n] leaves the last row half weighted:
    sum_{r < n-1} row_r + 1/2 row_{n-1} (intersection.h:38-39, forward_2d.hpp:63-75)."""
    n = 8
    img = rng(3).random((n, n))
    g = _axis_geom(n, np.pi)
    N, e, v, _ = oracle.Fan(g).ray(0, 3)
    assert N == n + 1
    np.testing.assert_allclose(e, [n / 2, -1], atol=1e-12)
    sino = oracle.fwd2d(g, img)
    cols = 0.5 * (img[:, n // 2 - 1] + img[:, n // 2])
    expect = cols[:-1].sum() + 0.5 * cols[-1]
    assert sino[0, 3] == pytest.approx(expect, rel=1e-12)


def test_horizontal_ray_closed_form(oracle):
    """theta = pi/2: source on +x, the centre ray runs along y = n/2 towards -x."""
    n = 8
    img = rng(4).random((n, n))
    g = _axis_geom(n, np.pi / 2)
    sino = oracle.fwd2d(g, img)
    expect = 0.5 * (img[n // 2 - 1, :] + img[n // 2, :]).sum()
    assert sino[0, 3] == pytest.approx(expect, rel=1e-12)


def test_missed_rays_are_zero(oracle):
    """A detector far wider than the image: outer cells miss the clip box and read 0."""
    g = fan(8, 61, 4, 16.0, 24.0)
    sino = oracle.fwd2d(g, np.ones((8, 8)))
    f = oracle.Fan(g)
    pl = f.plan()
    for a in range(4):
        for k in (0, 1, 59, 60):
            assert f.ray(a, k, pl)[0] == 0
            assert sino[a, k] == 0.0
    assert (sino[:, 28:33] > 0).all()


def test_forward_matches_analytic_ellipses(oracle):
    """Riemann sum with world step length ~ exact line integrals of the ellipse phantom
    (pins the scale convention: sum times hypot(v * spacing)). The shapes are halved so that
    none crosses the image border, which the analytic integral would count and the pixel
    grid truncates; the remaining error is point sampling of the edges, O(1/n)."""
    n = 256
    g = fan(n, 363, 24, 2.0 * n, 3.0 * n)
    p = phantom.ellipse_params(n, seed=5).copy()
    p[:, :4] *= 0.5
    img = phantom.ellipses_2d(n, params=p)
    sino = oracle.fwd2d(g, img)
    vec = np.array(FanGeometry(DSD=g.DSD, DSO=g.DSO, angles=g.angles, vol_shape=g.vol_shape,
                               det_shape=g.det_shape).to_vec())
    k = np.arange(g.det_shape) - (g.det_shape - 1) / 2
    src = vec[:, None, 0:2]
    cell = vec[:, None, 2:4] + k[None, :, None] * vec[:, None, 4:6]
    ana = phantom.ellipse_line_integrals(p, np.broadcast_to(src, cell.shape), cell - src)
    rel = np.linalg.norm(sino - ana) / np.linalg.norm(ana)
    assert rel < 0.015, rel
    assert sino.sum() / ana.sum() == pytest.approx(1.0, abs=2e-3)


# ------------------------------------------------------------------ adjoint --------------

def _system_matrix(oracle, g):
    ny, nx = int(g.vol_shape[0]), int(g.vol_shape[1])
    cols = []
    for j in range(nx * ny):
        e = np.zeros(nx * ny)
        e[j] = 1.0
        cols.append(oracle.fwd2d(g, e.reshape(ny, nx)).ravel())
    return np.stack(cols, axis=1)


def test_adjoint_is_transpose_of_assembled_matrix(oracle):
    g = tiny()
    M = _system_matrix(oracle, g)
    y = rng(6).random((g.nangles, g.det_shape))
    at = oracle.adj2d(g, y)
    np.testing.assert_allclose(at.ravel(), M.T @ y.ravel(), rtol=1e-12, atol=1e-12)


def test_dot_product_c1(oracle):
    g = c1()
    x = rng(7).random((64, 64))
    y = rng(8).random((360, 91))
    lhs = np.vdot(oracle.fwd2d(g, x), y)
    rhs = np.vdot(x, oracle.adj2d(g, y))
    assert abs(lhs - rhs) / abs(lhs) < 1e-12


def test_forward_linearity_and_positivity(oracle):
    g = c1()
    a = rng(9).random((64, 64))
    b = rng(10).random((64, 64))
    fa, fb = oracle.fwd2d(g, a), oracle.fwd2d(g, b)
    np.testing.assert_allclose(oracle.fwd2d(g, 2 * a - 3 * b), 2 * fa - 3 * fb, atol=1e-10)
    assert (fa >= 0).all()


# ------------------------------------------------------------------ backprojection (a7) --

@pytest.mark.parametrize("spacing,du", [((1.0, 1.0), 1.0), ((2.0, 0.5), 1.5)])
def test_backprojection_matches_vector_form(oracle, spacing, du):
    """Each pixel (at the reference's position ((i - n/2) s + o + 0.5), backward_2d.hpp:58-59)
    is projected from the source onto the detector line given by to_vec; the sample is
    weighted by DSD over its depth along the central ray, and the sum is divided by du."""
    g = fan(12, 21, 9, 40.0, 70.0, du=du, spacing=spacing, offset=(0.4, -0.6))
    g.angles = rng(11).uniform(0, 2 * np.pi, 9)
    sino = rng(12).random((9, 21))
    ours = oracle.bp2d(g, sino)
    vec = np.array(FanGeometry(DSD=g.DSD, DSO=g.DSO, angles=g.angles, vol_shape=g.vol_shape,
                               det_shape=g.det_shape).to_vec())
    ny, nx = 12, 12
    sy, sx = spacing
    oy, ox = 0.4, -0.6
    xs = (np.arange(nx) - nx / 2) * sx + ox + 0.5
    ys = (np.arange(ny) - ny / 2) * sy + oy + 0.5
    X, Y = np.meshgrid(xs, ys)
    P = np.stack([X, Y], -1)
    expect = np.zeros((ny, nx))
    for a in range(9):
        s, d, u = vec[a, 0:2], vec[a, 2:4], vec[a, 4:6]
        w = P - s
        # s + t w = d + q u  ->  [w, -u] [t, q]^T = d - s
        det = w[..., 0] * (-u[1]) - w[..., 1] * (-u[0])
        rhs = d - s
        q = (w[..., 0] * rhs[1] - w[..., 1] * rhs[0]) / det
        ucell = q / du + 21 / 2
        axis = (d - s) / np.linalg.norm(d - s)
        weight = g.DSD[a] / (w @ axis)
        row = np.concatenate([[0.0], sino[a], [0.0]])
        val = np.interp(ucell - 0.5, np.arange(-1, 22), row, left=0.0, right=0.0)
        expect += weight * val
    expect /= du
    np.testing.assert_allclose(ours, expect, rtol=1e-11, atol=1e-11)


def test_backprojection_accumulates(oracle):
    g = c1()
    y = rng(13).random((360, 91))
    base = rng(14).random((64, 64))
    np.testing.assert_allclose(oracle.bp2d(g, y, img=base), base + oracle.bp2d(g, y),
                               rtol=1e-13)


# ------------------------------------------------------------------ FP32 port ------------

def test_fp32_port_close_to_fp64(oracle):
    g = c1()
    img = phantom.ellipses_2d(64, seed=0)
    f64 = oracle.fwd2d(g, img)
    f32 = oracle.fwd2d_f32(g, img)
    assert np.linalg.norm(f32 - f64) / np.linalg.norm(f64) < 1e-5
    b64 = oracle.bp2d(g, f64)
    b32 = oracle.bp2d_f32(g, f64.astype(np.float32))
    assert np.linalg.norm(b32 - b64) / np.linalg.norm(b64) < 1e-5


# ------------------------------------------------------------------ golden fixtures ------

def test_golden_fixture_2d(oracle):
    import os
    d = np.load(os.path.join(os.path.dirname(__file__), "golden", "fan2d_tiny.npz"))
    g = tiny()
    np.testing.assert_array_equal(d["angles"], g.angles)
    np.testing.assert_allclose(oracle.fwd2d(g, d["image"]), d["sino"], rtol=1e-13, atol=1e-13)
    np.testing.assert_allclose(oracle.bp2d(g, d["sino_in"]), d["bp"], rtol=1e-13, atol=1e-13)
    np.testing.assert_allclose(oracle.adj2d(g, d["sino_in"]), d["adj"], rtol=1e-13, atol=1e-13)


# ------------------------------------------------------- vector geometry ----

def _views_from_circular(g):
    """ASTRA-like views of a circular record, built independently of the oracle: source at
    R(det_offset, DSO), detector centre at R(0, DSO - DSD), cell step R(du, 0), with R the
    reference's clockwise rotation (rotation.hpp:60) and the det_offset source shift of
    plan_2d.hpp:76-80."""
    out = []
    for th, dso, dsd in zip(g.angles, g.DSO, g.DSD):
        c, s = np.cos(th), np.sin(th)
        R = np.array([[c, s], [-s, c]])
        out.append(np.concatenate([R @ [g.det_offset, dso], R @ [0.0, dso - dsd],
                                   R @ [g.det_spacing, 0.0]]))
    return np.array(out)


def _with_views(g, views):
    r = SimpleNamespace(**vars(g))
    r.vecs = np.asarray(views, float)
    return r


def _general_views(seed, A, D, dso=60.0, dsd=100.0):
    """Views with a tilted, shifted detector and a per-view cell spacing."""
    r = rng(seed)
    out = []
    for th in r.uniform(0, 2 * np.pi, A):
        c, s = np.cos(th), np.sin(th)
        R = np.array([[c, s], [-s, c]])
        tilt = r.uniform(-0.35, 0.35)
        du = r.uniform(0.7, 1.4)
        src = R @ [r.uniform(-3, 3), dso]
        det = R @ [r.uniform(-4, 4), dso - dsd]
        u = R @ (du * np.array([np.cos(tilt), np.sin(tilt)]))
        out.append(np.concatenate([src, det, u]))
    return np.array(out)


def test_vector_views_reproduce_circular_plan_and_operators(oracle):
    g = fan(20, 31, 12, 50.0, 85.0, du=1.3, det_offset=0.7, spacing=(1.1, 0.9),
            offset=(0.3, -0.5))
    g.DSO = np.linspace(48.0, 55.0, 12)
    g.DSD = g.DSO + np.linspace(30.0, 40.0, 12)
    v = _with_views(g, _views_from_circular(g))
    for x, y in zip(oracle.Fan(g).plan(), oracle.Fan(v).plan()):
        np.testing.assert_allclose(y, x, rtol=1e-12, atol=1e-12)
    img = rng(20).random((20, 20))
    np.testing.assert_allclose(oracle.fwd2d(v, img), oracle.fwd2d(g, img), rtol=1e-11,
                               atol=1e-11)
    y = rng(21).random((12, 31))
    np.testing.assert_allclose(oracle.adj2d(v, y), oracle.adj2d(g, y), rtol=1e-11, atol=1e-11)
    # the circular backprojector ignores det_offset (backward_2d.hpp:58-99), the views honour it
    g.det_offset = 0.0
    v = _with_views(g, _views_from_circular(g))
    for wpow in (1, 2):
        np.testing.assert_allclose(oracle.bp2d(v, y, wpow=wpow), oracle.bp2d(g, y, wpow=wpow),
                                   rtol=1e-11, atol=1e-11)


@pytest.mark.parametrize("wpow", [1, 2])
def test_vector_backprojection_matches_line_intersection(oracle, wpow):
    """Each pixel (the reference's position, backward_2d.hpp:58-59) is joined to the source and
    the line cut with the detector line d + q u; the sample at cell q + (D - 1)/2 is weighted by
    the magnification t (pixel at s + (x - s), detector point at s + t (x - s)) to the power
    wpow and divided by the view's cell spacing |u|."""
    D, A = 23, 7
    g = fan(14, D, A, 60.0, 100.0, spacing=(1.2, 0.8), offset=(0.5, -0.4), ny=10)
    views = _general_views(22, A, D)
    v = _with_views(g, views)
    sino = rng(23).random((A, D))
    ours = oracle.bp2d(v, sino, wpow=wpow)
    xs = (np.arange(14) - 7) * 0.8 - 0.4 + 0.5
    ys = (np.arange(10) - 5) * 1.2 + 0.5 + 0.5
    X, Y = np.meshgrid(xs, ys)
    expect = np.zeros((10, 14))
    for a in range(A):
        s, d, u = views[a, 0:2], views[a, 2:4], views[a, 4:6]
        wx, wy = X - s[0], Y - s[1]
        rhs = d - s
        det = wx * (-u[1]) + wy * u[0]
        t = (rhs[0] * (-u[1]) + rhs[1] * u[0]) / det
        q = (wx * rhs[1] - wy * rhs[0]) / det
        row = np.concatenate([[0.0], sino[a], [0.0]])
        val = np.interp(q + (D - 1) / 2, np.arange(-1, D + 1), row, left=0.0, right=0.0)
        expect += t ** wpow * val / np.linalg.norm(u)
    np.testing.assert_allclose(ours, expect, rtol=1e-11, atol=1e-11)


def test_vector_forward_matches_analytic_ellipses(oracle):
    """Rays from the source through every cell centre d + (k - (D - 1)/2) u of general views
    against exact line integrals of the (halved) ellipse phantom, as for circular views."""
    n, D, A = 192, 301, 16
    g = fan(n, D, A, 400.0, 650.0)
    views = _general_views(24, A, D, dso=400.0, dsd=650.0)
    v = _with_views(g, views)
    p = phantom.ellipse_params(n, seed=6).copy()
    p[:, :4] *= 0.5
    sino = oracle.fwd2d(v, phantom.ellipses_2d(n, params=p))
    k = np.arange(D) - (D - 1) / 2
    src = views[:, None, 0:2]
    cell = views[:, None, 2:4] + k[None, :, None] * views[:, None, 4:6]
    ana = phantom.ellipse_line_integrals(p, np.broadcast_to(src, cell.shape), cell - src)
    rel = np.linalg.norm(sino - ana) / np.linalg.norm(ana)
    assert rel < 0.02, rel
    assert sino.sum() / ana.sum() == pytest.approx(1.0, abs=3e-3)


def test_vector_dot_product(oracle):
    g = fan(24, 35, 10, 60.0, 100.0)
    v = _with_views(g, _general_views(25, 10, 35))
    x = rng(26).random((24, 24))
    y = rng(27).random((10, 35))
    lhs = np.vdot(oracle.fwd2d(v, x), y)
    rhs = np.vdot(x, oracle.adj2d(v, y))
    assert abs(lhs - rhs) / abs(lhs) < 1e-12


# ---------------------------------------------------------- exact detector geometry -------

def _exact(rec, cor=0.0, rot=0.0):
    from types import SimpleNamespace
    r = SimpleNamespace(**vars(rec))
    r.exact_geometry, r.COR, r.det_rotation = True, float(cor), float(rot)
    return r


@pytest.mark.parametrize("k", [3, -2])
def test_exact_offset_is_a_shifted_wider_detector(oracle, k):
    """An exact detector offset of k whole cells equals the centred detector with 2|k| more
    cells, read k + |k| cells further on: the rays are the same lines (footprint check of the
    offset's direction and size), in the forward and, for a sinogram that is zero outside the
    shifted window, in the backprojection."""
    D, Dw = 21, 21 + 2 * abs(k)
    base = fan(16, D, 10, 40.0, 64.0, det_offset=float(k), spacing=(1.0, 1.0))
    wide = fan(16, Dw, 10, 40.0, 64.0)
    img = np.random.default_rng(3).random((16, 16))
    so = oracle.fwd2d(_exact(base), img)
    sw = oracle.fwd2d(wide, img)
    j0 = k + abs(k)
    np.testing.assert_allclose(so, sw[:, j0:j0 + D], rtol=1e-9, atol=1e-9)
    y = np.random.default_rng(4).random((10, D))
    yw = np.zeros((10, Dw))
    yw[:, j0:j0 + D] = y
    np.testing.assert_allclose(oracle.bp2d(_exact(base), y), oracle.bp2d(wide, yw), rtol=1e-9,
                               atol=1e-9)
    # without the exact flag the reference's semantics hold: the forward moves the source
    assert not np.allclose(oracle.fwd2d(base, img), so)


def test_exact_cor_moves_the_rotation_centre(oracle):
    """For one projection, COR = c is the same as moving the image by c along the detector
    axis u = (cos t, -sin t) (the reference's clockwise rotation of (1, 0))."""
    t, c = 0.7, 1.6
    rec = fan(14, 19, 1, 30.0, 52.0, angles=[t])
    img = np.random.default_rng(5).random((14, 14))
    a = oracle.fwd2d(_exact(rec, cor=c), img)
    moved = fan(14, 19, 1, 30.0, 52.0, angles=[t], offset=(-c * np.sin(t), c * np.cos(t)))
    b = oracle.fwd2d(_exact(moved), img)
    np.testing.assert_allclose(a, b, rtol=1e-9, atol=1e-9)


def test_exact_rotation_turns_the_detector_about_its_centre(oracle):
    """With the detector turned by r, the ray of cell k passes through the point
    centre + (k - (D - 1)/2) du (cos r, sin r) before the rotation by the angle (a line check on
    the planned rays), and the centre cell's ray does not move."""
    r_ = 0.2
    rec = fan(12, 15, 4, 30.0, 50.0, du=1.3)
    ex = _exact(rec, rot=r_)
    f = oracle.Fan(ex)
    pl = f.plan()
    for a, t in enumerate(rec.angles):
        c, s = np.cos(t), np.sin(t)
        R = np.array([[c, s], [-s, c]])
        src = R @ np.array([0.0, 30.0])
        for k in (0, 7, 14):
            p = R @ (np.array([0.0, -20.0]) + (k - 7) * 1.3 * np.array([np.cos(r_), np.sin(r_)]))
            n, e, v, sc = f.ray(a, k, pl)
            # the ray (texel frame = world + n/2 per axis at spacing 1, texel i centred at
            # i + 0.5) is the line src -> p
            ew = e - 12 / 2
            d = p - src

            def cross(a, b):
                return a[0] * b[1] - a[1] * b[0]
            assert abs(cross(d, ew - src)) < 1e-9 * np.linalg.norm(d) ** 2
            assert abs(cross(d, v)) < 1e-12 * np.linalg.norm(d) * np.linalg.norm(v)
    img = np.random.default_rng(6).random((12, 12))
    np.testing.assert_allclose(oracle.fwd2d(ex, img)[:, 7], oracle.fwd2d(_exact(rec), img)[:, 7],
                               rtol=1e-12)


def test_exact_geometry_adjoint_is_transpose(oracle):
    rec = _exact(fan(6, 9, 5, 20.0, 33.0, det_offset=0.7, du=1.2), cor=0.9, rot=0.1)
    cols = []
    for j in range(36):
        e = np.zeros(36)
        e[j] = 1.0
        cols.append(oracle.fwd2d(rec, e.reshape(6, 6)).ravel())
    M = np.stack(cols, axis=1)
    y = np.random.default_rng(7).random((5, 9))
    np.testing.assert_allclose(oracle.adj2d(rec, y).ravel(), M.T @ y.ravel(), rtol=1e-12,
                               atol=1e-12)


# ------------------------------------------------- transpose of the backprojection --------

def _bp_matrix(oracle, g, wpow):
    """The backprojection as a matrix (image pixels x sinogram cells), one cell at a time."""
    A, D = g.nangles, int(g.det_shape)
    ny, nx = (int(v) for v in g.vol_shape)
    M = np.zeros((ny * nx, A * D))
    for j in range(A * D):
        e = np.zeros(A * D)
        e[j] = 1.0
        M[:, j] = oracle.bp2d(g, e.reshape(A, D), wpow=wpow).ravel()
    return M


@pytest.mark.parametrize("wpow", [1, 2])
@pytest.mark.parametrize("views", [False, True])
def test_backprojection_transpose_is_assembled_matrix_transpose(oracle, wpow, views):
    g = fan(8, 9, 6, 16.0, 24.0, du=1.2, spacing=(1.1, 0.9), offset=(0.3, -0.2))
    if views:
        g = _with_views(g, _general_views(21, 6, 9, dso=16.0, dsd=24.0))
    M = _bp_matrix(oracle, g, wpow)
    img = rng(22).random((8, 8))
    bt = oracle.bp2d_T(g, img, wpow=wpow)
    np.testing.assert_allclose(bt.ravel(), M.T @ img.ravel(), rtol=1e-12, atol=1e-12)
    # a projection subset is the same rows of the full transpose
    part = oracle.bp2d_T(g, img, angles=(2, 4), wpow=wpow)
    np.testing.assert_allclose(part[2:4], bt[2:4], rtol=0, atol=0)
    assert not part[:2].any() and not part[4:].any()


def test_backprojection_transpose_dot_product_c1(oracle):
    g = c1()
    s = rng(23).random((360, 91))
    v = rng(24).random((64, 64))
    lhs = np.vdot(oracle.bp2d(g, s), v)
    rhs = np.vdot(s, oracle.bp2d_T(g, v))
    assert abs(lhs - rhs) / abs(lhs) < 1e-12
