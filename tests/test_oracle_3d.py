"""Pins the FP64 oracle's 3D routines without restating its formulas: rotations against
scipy's Euler conventions, the forward geometry against an independent vector construction,
an axis-aligned ray whose samples fall on voxel centres (closed form), analytic ellipsoid line
integrals, the backprojectors against line-plane / linear-solve restatements with scipy
interpolation, and the adjoint against an explicitly assembled system matrix."""
import numpy as np
import pytest
from scipy import ndimage
from scipy.spatial.transform import Rotation

from curadon_amd import phantom
from geoms import cone, helix_views as _helix, with_views as _with_views


def rng(seed):
    return np.random.default_rng(seed)


def test_rotations_match_scipy(oracle):
    g = rng(0)
    for _ in range(20):
        v = g.normal(size=3)
        a, b, c = g.uniform(-np.pi, np.pi, 3)
        # rotate_yzy = Ry(phi) Rz(theta) Ry(psi); rotate_roll_pitch_yaw = Rz(roll) Ry(pitch) Rx(yaw)
        np.testing.assert_allclose(oracle.rotate_yzy(v, a, b, c),
                                   Rotation.from_euler("YZY", [a, b, c]).apply(v), atol=1e-13)
        np.testing.assert_allclose(oracle.rotate_rpy(v, a, b, c),
                                   Rotation.from_euler("ZYX", [a, b, c]).apply(v), atol=1e-13)


def _fwd_vector_form(g, a):
    """Forward frame built from scratch: panel pixel (iu, iv) at (-du U/2 + du/2 + iu du, ...,
    0), detector rotation, +DOD, trajectory rotation, then into voxel coordinates."""
    V, U = g.det_shape
    dv, du = g.det_spacing
    sp = g.vol_spacing[::-1]
    off = g.vol_offset[::-1]
    n = g.vol_shape[::-1].astype(float)
    pitch, roll, yaw = g.det_rotation
    phi, theta, psi = g.angles[:, a]
    Rd = Rotation.from_euler("ZYX", [roll, pitch, yaw])
    Rt = Rotation.from_euler("YZY", [phi, theta, psi])
    dod = g.DSD - g.DSO

    def pix(iu, iv):
        p = np.array([-du * U / 2 + du / 2 + iu * du, -dv * V / 2 + dv / 2 + iv * dv, 0.0])
        return Rd.apply(p) + [0, 0, dod]

    def to_vox(w):
        x = (Rt.apply(w) - off + n * sp / 2 - sp / 2) / sp
        return x + [g.COR * np.sin(phi) / sp[0], 0, g.COR * np.cos(phi) / sp[2]]

    o = to_vox(pix(0, 0))
    return o, to_vox(pix(1, 0)) - o, to_vox(pix(0, 1)) - o, to_vox(np.array([0, 0, -g.DSO]))


def test_forward_geometry_matches_vector_form(oracle):
    g = cone(10, 9, 7, 5, 40.0, 65.0, du=(1.3, 0.8), vol_spacing=(0.9, 1.1, 1.2),
             vol_offset=(0.3, -0.4, 0.5), det_rotation=(0.05, -0.1, 0.08), cor=0.7,
             angles=rng(1).uniform(-1, 1, (3, 5)) * np.array([[3.0], [0.4], [0.4]]), nz=8, ny=9)
    org, du, dv, src = oracle.Cone(g).plan_fwd()
    for a in range(5):
        o, u, v, s = _fwd_vector_form(g, a)
        np.testing.assert_allclose(org[a], o, atol=1e-11)
        np.testing.assert_allclose(du[a], u, atol=1e-11)
        np.testing.assert_allclose(dv[a], v, atol=1e-11)
        np.testing.assert_allclose(src[a], s, atol=1e-11)


def test_axis_ray_closed_form(oracle):
    """phi = 0, DSO = 16.5, DSD = 32, n = 8, odd panel: the centre ray runs along z at
    x = y = 3.5 from z = -1 to z = 9 in exactly 10 steps of one voxel (all values exact in
    binary), so the unscaled sum is sum_z of the mean of the four voxels around the axis."""
    g = cone(8, 7, 7, 1, 16.5, 32.0, angles=[0.0])
    vol = rng(2).random((8, 8, 8))
    nst, s, d, t0, dt = oracle.Cone(g).ray(0, 3, 3)
    assert nst == 10
    np.testing.assert_allclose(d, [0, 0, 32.0], atol=0)
    assert dt * 32.0 == 1.0
    sino = oracle.fwd3d(g, vol)
    expect = 0.25 * (vol[:, 3, 3] + vol[:, 3, 4] + vol[:, 4, 3] + vol[:, 4, 4]).sum()
    assert sino[0, 3, 3] == pytest.approx(expect, rel=1e-13)


def test_missed_rays_are_zero(oracle):
    g = cone(8, 41, 41, 2, 16.5, 32.0)
    sino = oracle.fwd3d(g, np.ones((8, 8, 8)))
    c = oracle.Cone(g)
    pl = c.plan_fwd()
    for a in range(2):
        assert c.ray(a, 0, 0, pl)[0] == -1
        assert sino[a, 0, 0] == 0.0 and sino[a, 40, 40] == 0.0
    assert sino[:, 20, 20].min() > 0


def test_forward_times_step_matches_analytic_ellipsoids(oracle):
    """sum of samples x step length ~ exact line integrals of ellipsoids that lie inside."""
    n = 48
    g = cone(n, 72, 72, 6, 96.0, 144.0)
    p = phantom.ellipsoid_params(n, seed=3).copy()
    p[:, :3] *= 0.4  # centres within 0.16 n, semi-axes <= 0.31 n: every shape inside
    vol = phantom.ellipsoids_3d(n, params=p, dtype=np.float64)
    sino = oracle.fwd3d(g, vol)
    c = oracle.Cone(g)
    pl = c.plan_fwd()
    got, ana = [], []
    for a in range(6):
        for iv in range(4, 72, 5):
            for iu in range(4, 72, 5):
                nst, s, d, t0, dt = c.ray(a, iu, iv, pl)
                if nst < 0:
                    continue
                got.append(sino[a, iv, iu] * dt * np.linalg.norm(d))
                src_w = s - (n / 2 - 0.5)     # voxel -> world, spacing 1
                ana.append(phantom.ellipsoid_line_integrals(p, src_w, d))
    got, ana = np.array(got), np.array(ana)
    assert np.linalg.norm(got - ana) / np.linalg.norm(ana) < 0.04
    assert got.sum() / ana.sum() == pytest.approx(1.0, abs=5e-3)


def _assemble(oracle, g):
    nz, ny, nx = (int(v) for v in g.vol_shape)
    cols = []
    for j in range(nx * ny * nz):
        e = np.zeros(nx * ny * nz)
        e[j] = 1.0
        cols.append(oracle.fwd3d(g, e.reshape(nz, ny, nx)).ravel())
    return np.stack(cols, axis=1)


def test_adjoint_is_transpose_of_assembled_matrix(oracle):
    g = cone(5, 4, 3, 3, 12.0, 20.0, nz=4, ny=6)
    M = _assemble(oracle, g)
    y = rng(4).random((3, 3, 4))
    np.testing.assert_allclose(oracle.adj3d(g, y).ravel(), M.T @ y.ravel(), rtol=1e-12,
                               atol=1e-12)


def test_dot_product_3d(oracle):
    g = cone(16, 20, 18, 12, 40.0, 60.0)
    x = rng(5).random((16, 16, 16))
    y = rng(6).random((12, 18, 20))
    lhs = np.vdot(oracle.fwd3d(g, x), y)
    rhs = np.vdot(x, oracle.adj3d(g, y))
    assert abs(lhs - rhs) / abs(lhs) < 1e-12


def _voxel_world(g):
    nz, ny, nx = (int(v) for v in g.vol_shape)
    sp = g.vol_spacing[::-1]
    off = g.vol_offset[::-1]
    n = np.array([nx, ny, nz], float)
    base = -n * sp / 2 + sp / 2 + off
    Z, Y, X = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    return base + np.stack([X, Y, Z], -1) * sp


@pytest.mark.parametrize("dsd", [40.0, 64.0])
def test_reference_backprojection_as_line_plane_intersection(oracle, dsd):
    """Reference mode: voxel positions rotated by Ryzy (not its inverse), then the line from
    the source (0, 0, -DSO) through each voxel is cut with the plane z = DSO - DSD; (x, y) of
    the cut plus floor(U/2), floor(V/2) is the texel coordinate (backward_3d.hpp:79-94)."""
    g = cone(10, 14, 12, 5, 40.0, dsd, nz=9, ny=11,
             angles=np.vstack([rng(7).uniform(0, 2 * np.pi, 5), np.zeros(5), np.zeros(5)]))
    sino = rng(8).random((5, 12, 14))
    ours = oracle.bp3d(g, sino)
    P = _voxel_world(g)
    plane = g.DSO - g.DSD
    s = np.array([0.0, 0.0, -g.DSO])
    expect = np.zeros(P.shape[:-1])
    for a in range(5):
        Q = Rotation.from_euler("YZY", g.angles[:, a]).apply(P.reshape(-1, 3)).reshape(P.shape)
        lam = (plane - s[2]) / (Q[..., 2] - s[2])
        cut = s + lam[..., None] * (Q - s)
        u = cut[..., 0] + 14 // 2
        v = cut[..., 1] + 12 // 2
        expect += ndimage.map_coordinates(sino[a], [v - 0.5, u - 0.5], order=1,
                                          mode="grid-constant").reshape(u.shape)
    np.testing.assert_allclose(ours, expect, rtol=1e-11, atol=1e-11)


def test_matched_backprojection_as_linear_solve(oracle):
    """Matched mode: each voxel centre P solves P = src + t (org + iu du + iv dv - src) in the
    forward frame; the panel is sampled at pixel coordinate (iu, iv); voxels behind the source
    get nothing."""
    g = cone(9, 13, 11, 4, 30.0, 55.0, nz=8, ny=10, det_rotation=(0.05, 0.02, -0.03),
             angles=rng(9).uniform(0, 2 * np.pi, 4))
    sino = rng(10).random((4, 11, 13))
    ours = oracle.bp3d(g, sino, matched=True)
    org, du, dv, src = oracle.Cone(g).plan_fwd()
    nz, ny, nx = (int(v) for v in g.vol_shape)
    Z, Y, X = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    P = np.stack([X, Y, Z], -1).astype(float)
    expect = np.zeros(P.shape[:-1])
    for a in range(4):
        M = np.stack([org[a] - src[a], du[a], dv[a]], axis=1)
        x = np.linalg.solve(M, (P - src[a]).reshape(-1, 3).T).T.reshape(P.shape)
        iu, iv = x[..., 1] / x[..., 0], x[..., 2] / x[..., 0]
        val = ndimage.map_coordinates(sino[a], [iv, iu], order=1, mode="grid-constant")
        expect += np.where(x[..., 0] > 0, val.reshape(iu.shape), 0.0)
    np.testing.assert_allclose(ours, expect, rtol=1e-10, atol=1e-10)


def test_matched_backprojection_hits_forward_footprint(oracle):
    """A one-pixel sinogram backprojected with the matched geometry lights up voxels along that
    pixel's forward ray: the forward projection of the backprojected volume peaks at the pixel."""
    g = cone(16, 15, 15, 1, 48.0, 72.0, angles=[0.7])
    y = np.zeros((1, 15, 15))
    y[0, 6, 9] = 1.0
    back = oracle.bp3d(g, y, matched=True)
    reproj = oracle.fwd3d(g, back)
    assert np.unravel_index(np.argmax(reproj[0]), (15, 15)) == (6, 9)


def test_backprojection_accumulates(oracle):
    g = cone(6, 8, 7, 3, 20.0, 30.0)
    y = rng(11).random((3, 7, 8))
    base = rng(12).random((6, 6, 6))
    np.testing.assert_allclose(oracle.bp3d(g, y, vol=base), base + oracle.bp3d(g, y), rtol=1e-13)


def test_golden_fixture_3d(oracle):
    import os
    d = np.load(os.path.join(os.path.dirname(__file__), "golden", "cone3d_tiny.npz"))
    g = cone(8, 6, 5, 8, 16.0, 28.0)
    np.testing.assert_array_equal(d["angles"], g.angles)
    np.testing.assert_allclose(oracle.fwd3d(g, d["volume"]), d["sino"], rtol=1e-13, atol=1e-13)
    np.testing.assert_allclose(oracle.bp3d(g, d["sino_in"]), d["bp"], rtol=1e-13, atol=1e-13)
    np.testing.assert_allclose(oracle.bp3d(g, d["sino_in"], matched=True), d["bp_matched"],
                               rtol=1e-13, atol=1e-13)
    np.testing.assert_allclose(oracle.adj3d(g, d["sino_in"]), d["adj"], rtol=1e-13, atol=1e-13)


# ------------------------------------------------------------------ views (vector form) ----

def _views_from_scratch(g):
    """World views (source, panel centre, u-step, v-step) of an Euler-angle geometry built with
    scipy's rotations, independently of the oracle and of the package's to_vec."""
    dv, du = g.det_spacing
    pitch, roll, yaw = g.det_rotation
    Rd = Rotation.from_euler("ZYX", [roll, pitch, yaw])
    dod = g.DSD - g.DSO
    out = []
    for a in range(g.angles.shape[1]):
        phi, theta, psi = g.angles[:, a]
        Rt = Rotation.from_euler("YZY", [phi, theta, psi])
        cor = g.COR * np.array([np.sin(phi), 0.0, np.cos(phi)])
        # the panel centre: pixel (0, 0) at (-du U/2 + du/2, -dv V/2 + dv/2) plus (U-1)/2 and
        # (V-1)/2 steps is the panel frame's origin, which sits at DOD along z
        c = Rt.apply([0.0, 0.0, dod]) + cor
        u = Rt.apply(Rd.apply([du, 0.0, 0.0]))
        v = Rt.apply(Rd.apply([0.0, dv, 0.0]))
        s = Rt.apply([0.0, 0.0, -g.DSO]) + cor
        out.append(np.concatenate([s, c, u, v]))
    return np.array(out)


def test_views_of_euler_geometry_give_its_forward_frame(oracle):
    """Views built from an Euler-angle geometry (tilted orbit, rotated and offset panel,
    anisotropic voxels, COR) reproduce the Euler set-up's frames, so the vector path is pinned
    to the Euler path, which is pinned against scipy's rotations above."""
    g = cone(10, 9, 7, 5, 40.0, 65.0, du=(1.3, 0.8), vol_spacing=(0.9, 1.1, 1.2),
             vol_offset=(0.3, -0.4, 0.5), det_rotation=(0.05, -0.1, 0.08), cor=0.7,
             angles=rng(1).uniform(-1, 1, (3, 5)) * np.array([[3.0], [0.4], [0.4]]), nz=8, ny=9)
    gv = _with_views(g, _views_from_scratch(g))
    for a, b in zip(oracle.Cone(g).plan_fwd(), oracle.Cone(gv).plan_fwd()):
        np.testing.assert_allclose(a, b, atol=1e-11)
    vol = rng(4).random((8, 9, 10))
    np.testing.assert_allclose(oracle.fwd3d(gv, vol), oracle.fwd3d(g, vol), rtol=1e-10,
                               atol=1e-10)


def test_view_rays_join_source_and_pixel(oracle):
    """Line-plane check of the vector path: every ray of a helical, tilted-panel geometry
    starts at the view's source and reaches the world point centre + (iu - (U-1)/2) u +
    (iv - (V-1)/2) v at parameter 1 (voxel -> world: p sp - (n sp / 2 - sp / 2) + offset)."""
    g = cone(12, 9, 7, 6, 30.0, 50.0, vol_spacing=(1.1, 0.9, 1.2), vol_offset=(0.4, -0.2, 0.3),
             nz=10, ny=11)
    vecs = _helix(6, 12, 30.0, 8.0, 20.0, 9, 7, du=1.3, dv=0.8, tilt=0.2)
    gv = _with_views(g, vecs)
    c = oracle.Cone(gv)
    pl = c.plan_fwd()
    sp = g.vol_spacing[::-1]
    off = g.vol_offset[::-1]
    n = g.vol_shape[::-1].astype(float)

    def to_world(p):
        return p * sp - (n * sp / 2 - sp / 2) + off
    for a in range(6):
        for iu, iv in [(0, 0), (8, 6), (3, 5), (4, 3)]:
            org, du, dv, src = (x[a] for x in pl)
            pix = org + iu * du + iv * dv
            want = vecs[a, 3:6] + (iu - 4) * vecs[a, 6:9] + (iv - 3) * vecs[a, 9:12]
            np.testing.assert_allclose(to_world(pix), want, atol=1e-12)
            np.testing.assert_allclose(to_world(src), vecs[a, 0:3], atol=1e-12)


def test_views_forward_matches_analytic_ellipsoids(oracle):
    """Forward projection along a helix with a tilted panel: sum of samples times step length
    against exact ellipsoid line integrals (same bar as the circular case above)."""
    n = 48
    g = cone(n, 72, 72, 6, 96.0, 144.0)
    vecs = _helix(6, n, 96.0, 30.0, 48.0, 72, 72, tilt=0.15)
    gv = _with_views(g, vecs)
    p = phantom.ellipsoid_params(n, seed=3).copy()
    p[:, :3] *= 0.4
    vol = phantom.ellipsoids_3d(n, params=p, dtype=np.float64)
    sino = oracle.fwd3d(gv, vol)
    c = oracle.Cone(gv)
    pl = c.plan_fwd()
    got, ana = [], []
    for a in range(6):
        for iv in range(4, 72, 5):
            for iu in range(4, 72, 5):
                nst, s, d, t0, dt = c.ray(a, iu, iv, pl)
                if nst < 0:
                    continue
                got.append(sino[a, iv, iu] * dt * np.linalg.norm(d))
                ana.append(phantom.ellipsoid_line_integrals(p, s - (n / 2 - 0.5), d))
    got, ana = np.array(got), np.array(ana)
    assert np.linalg.norm(got - ana) / np.linalg.norm(ana) < 0.04
    assert got.sum() / ana.sum() == pytest.approx(1.0, abs=5e-3)


def test_views_adjoint_and_matched_backprojection(oracle):
    """With views: the adjoint is the transpose of the assembled system matrix, the reference
    mode equals the matched backprojection, and the matched backprojection agrees with a
    linear solve per voxel (the restatement used for the Euler case)."""
    g = cone(4, 5, 4, 3, 12.0, 20.0, nz=3, ny=4)
    vecs = _helix(3, 4, 12.0, 3.0, 8.0, 5, 4, du=1.1, dv=0.9, tilt=0.1)
    gv = _with_views(g, vecs)
    M = _assemble(oracle, gv)
    y = rng(9).random((3, 4, 5))
    np.testing.assert_allclose(oracle.adj3d(gv, y).ravel(), M.T @ y.ravel(), rtol=1e-12,
                               atol=1e-12)
    np.testing.assert_array_equal(oracle.bp3d(gv, y), oracle.bp3d(gv, y, matched=True))
    with pytest.raises(ValueError):
        oracle.bp3d(gv, y, fdk=True)


def test_fp32_ports_follow_the_fp64_oracle(oracle):
    """The FP32 CPU ports timed as the CPU baseline compute the same operators: within FP32
    rounding (and the float-carried march's step count) of the FP64 oracle."""
    g = cone(16, 20, 18, 6, 40.0, 64.0)
    vol = rng(12).random((16, 16, 16))
    ref = oracle.fwd3d(g, vol)
    got = oracle.fwd3d_f32(g, vol)
    assert np.linalg.norm(got - ref) / np.linalg.norm(ref) < 1e-5
    y = rng(13).random((6, 18, 20))
    ref = oracle.bp3d(g, y)
    got = oracle.bp3d_f32(g, y)
    assert np.linalg.norm(got - ref) / np.linalg.norm(ref) < 1e-5


# ------------------------------------------------- transpose of the backprojection --------

def _bp_matrix3(oracle, g, **mode):
    """A 3D backprojection as a matrix (voxels x panel cells), one cell at a time."""
    A = g.angles.shape[1]
    V, U = (int(v) for v in g.det_shape)
    cols = []
    for j in range(A * V * U):
        e = np.zeros(A * V * U)
        e[j] = 1.0
        cols.append(oracle.bp3d(g, e.reshape(A, V, U), **mode).ravel())
    return np.stack(cols, axis=1)


@pytest.mark.parametrize("mode", [{}, {"matched": True}, {"fdk": True}])
def test_backprojection_transpose_is_assembled_matrix_transpose(oracle, mode):
    g = cone(5, 6, 4, 3, 12.0, 20.0, du=(1.1, 0.9), nz=4, ny=3, vol_spacing=(1.0, 1.2, 0.9),
             vol_offset=(0.3, -0.2, 0.1), det_offset=(0.4, -0.3))
    M = _bp_matrix3(oracle, g, **mode)
    x = rng(31).random((4, 3, 5))
    np.testing.assert_allclose(oracle.bp3d_T(g, x, **mode).ravel(), M.T @ x.ravel(),
                               rtol=1e-12, atol=1e-12)
    # a projection subset is the same projections of the full transpose
    full = oracle.bp3d_T(g, x, **mode)
    part = oracle.bp3d_T(g, x, angles=(1, 2), **mode)
    np.testing.assert_array_equal(part[1], full[1])
    assert not part[0].any() and not part[2].any()


def test_backprojection_transpose_with_views(oracle):
    g = cone(4, 5, 4, 3, 12.0, 20.0, nz=3, ny=4)
    gv = _with_views(g, _helix(3, 4, 12.0, 3.0, 8.0, 5, 4, du=1.1, dv=0.9, tilt=0.1))
    M = _bp_matrix3(oracle, gv)
    x = rng(32).random((3, 4, 4))
    np.testing.assert_allclose(oracle.bp3d_T(gv, x).ravel(), M.T @ x.ravel(), rtol=1e-12,
                               atol=1e-12)


def test_backprojection_transpose_dot_product_3d(oracle):
    g = cone(16, 20, 18, 12, 40.0, 60.0)
    s = rng(33).random((12, 18, 20))
    x = rng(34).random((16, 16, 16))
    for mode in ({}, {"matched": True}, {"fdk": True}):
        lhs = np.vdot(oracle.bp3d(g, s, **mode), x)
        rhs = np.vdot(s, oracle.bp3d_T(g, x, **mode))
        assert abs(lhs - rhs) / abs(lhs) < 1e-12
