"""GPU parity of the 3D cone-beam HIP kernels against the FP64 oracle, through the C-ABI.

Tolerances (FP32 storage and arithmetic vs FP64):
  forward                  rel-L2 <= 1e-5, max-abs <= 2e-5 * max|ref|
  backprojection (both)    rel-L2 <= 1e-5, max-abs <= 2e-5 * max|ref|
  adjoint                  rel-L2 <= 1e-5, max-abs <= 2e-5 * max|ref|
  dot test                 |<Ax,y> - <x,A^T y>| / |<Ax,y>| <= 1e-5
At the 256^3 workload the oracle runs on projection / slab subsets.
"""
import numpy as np
import pytest
import torch

import curadon_amd as cm
from curadon_amd import phantom
from geoms import as_cone_geometry, c3, cone

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


SMALL = dict(n=32, U=48, V=40, A=24, dso=64.0, dsd=96.0)


@pytest.fixture(scope="module")
def small(gpu):
    rec = cone(SMALL["n"], SMALL["U"], SMALL["V"], SMALL["A"], SMALL["dso"], SMALL["dsd"])
    return rec, as_cone_geometry(rec)


@pytest.fixture(scope="module")
def general(gpu):
    """Non-circular trajectory, rotated and offset panel, anisotropic voxels, COR shift."""
    r = np.random.default_rng(0)
    ang = np.vstack([r.uniform(0, 2 * np.pi, 10), r.uniform(-0.3, 0.3, 10),
                     r.uniform(-0.3, 0.3, 10)])
    rec = cone(24, 36, 30, 10, 60.0, 100.0, du=(1.2, 0.9), angles=ang,
               vol_spacing=(1.1, 0.9, 1.0), vol_offset=(0.5, -0.7, 0.3),
               det_offset=(1.5, -2.0), det_rotation=(0.04, -0.06, 0.03), cor=0.8, nz=20, ny=22)
    return rec, as_cone_geometry(rec)


def test_small_forward(oracle, small, gpu):
    rec, geom = small
    vol = phantom.ellipsoids_3d(32, seed=1, dtype=np.float64)
    ref = oracle.fwd3d(rec, vol)
    check_close(host(cm.forward(dev(vol, gpu), geom)), ref, "small fwd3d")


@pytest.mark.parametrize("spacing", [0.5, 0.12])
def test_forward_wide_beams(oracle, gpu, spacing):
    """Voxels smaller than the pixels' footprint: a 16 x 16-pixel workgroup's beam spans more
    voxels than one staged box holds at oblique angles (voxel 0.5) or everywhere (0.12), and
    those slabs take the global path. The forward still matches the oracle."""
    rec = cone(48, 40, 36, 12, 120.0, 180.0, vol_spacing=(spacing,) * 3)
    geom = as_cone_geometry(rec)
    vol = phantom.ellipsoids_3d(48, seed=3, dtype=np.float64)
    ref = oracle.fwd3d(rec, vol)
    out = host(cm.forward(dev(vol, gpu), geom))
    check_close(out, ref, f"wide-beam forward (voxel {spacing})")


def test_small_backward_reference(oracle, small, gpu):
    rec, geom = small
    y = np.random.default_rng(2).random((24, 40, 48))
    ref = oracle.bp3d(rec, y)
    check_close(host(cm.backward(dev(y, gpu), geom)), ref, "small bp3d reference")


def test_small_backward_matched(oracle, small, gpu):
    rec, geom = small
    y = np.random.default_rng(3).random((24, 40, 48))
    ref = oracle.bp3d(rec, y, matched=True)
    check_close(host(cm.backward(dev(y, gpu), geom, mode="matched")), ref, "small bp3d matched")


def test_small_adjoint_and_dot(oracle, small, gpu):
    rec, geom = small
    y = np.random.default_rng(4).random((24, 40, 48))
    ref = oracle.adj3d(rec, y)
    got = cm.adjoint(dev(y, gpu), geom)
    check_close(host(got), ref, "small adj3d")
    x = torch.rand(32, 32, 32, device=gpu, generator=torch.Generator(device=gpu).manual_seed(5))
    lhs = torch.vdot(cm.forward(x, geom).double().ravel(), dev(y, gpu).double().ravel())
    rhs = torch.vdot(x.double().ravel(), got.double().ravel())
    r = abs(float(lhs - rhs)) / abs(float(lhs))
    print(f"small dot test: {r:.3e}")
    assert r <= 1e-5


def test_general_geometry(oracle, general, gpu):
    rec, geom = general
    vol = np.random.default_rng(6).random((20, 22, 24))
    check_close(host(cm.forward(dev(vol, gpu), geom)), oracle.fwd3d(rec, vol), "general fwd3d")
    y = np.random.default_rng(7).random((10, 30, 36))
    check_close(host(cm.backward(dev(y, gpu), geom)), oracle.bp3d(rec, y), "general bp3d ref")
    check_close(host(cm.backward(dev(y, gpu), geom, mode="matched")),
                oracle.bp3d(rec, y, matched=True), "general bp3d matched")
    check_close(host(cm.adjoint(dev(y, gpu), geom)), oracle.adj3d(rec, y), "general adj3d")


def test_tilted_orbits_use_every_volume_copy(oracle, gpu):
    """The forward reads, per projection, the volume copy in which the axis closest to the
    detector's v direction is fastest. theta = 0 keeps v along y; theta = pi/2 turns it into the
    xz-plane, along x for phi near 0 and along z for phi near pi/2, so this plan stages all three
    copies. Batched, so that every copy is staged per image."""
    r = np.random.default_rng(11)
    phi = np.concatenate([np.linspace(0, 2 * np.pi, 4, endpoint=False),
                          r.uniform(-0.3, 0.3, 4), np.pi / 2 + r.uniform(-0.3, 0.3, 4)])
    theta = np.concatenate([np.zeros(4), np.full(8, np.pi / 2) + r.uniform(-0.2, 0.2, 8)])
    psi = r.uniform(-0.2, 0.2, 12)
    rec = cone(20, 30, 26, 12, 50.0, 80.0, angles=np.vstack([phi, theta, psi]), nz=18, ny=22)
    geom = as_cone_geometry(rec)
    vols = r.random((2, 18, 22, 20))
    got = host(cm.forward(dev(vols, gpu), geom))
    for i in range(2):
        check_close(got[i], oracle.fwd3d(rec, vols[i]), f"tilted fwd3d [{i}]")


def test_tiny_golden_3d(gpu):
    import os
    d = np.load(os.path.join(os.path.dirname(__file__), "golden", "cone3d_tiny.npz"))
    geom = as_cone_geometry(cone(8, 6, 5, 8, 16.0, 28.0))
    check_close(host(cm.forward(dev(d["volume"], gpu), geom)), d["sino"], "tiny fwd3d")
    check_close(host(cm.backward(dev(d["sino_in"], gpu), geom)), d["bp"], "tiny bp3d")
    check_close(host(cm.backward(dev(d["sino_in"], gpu), geom, mode="matched")),
                d["bp_matched"], "tiny bp3d matched")
    check_close(host(cm.adjoint(dev(d["sino_in"], gpu), geom)), d["adj"], "tiny adj3d")


def test_batched_equals_single_3d(small, gpu):
    rec, geom = small
    xs = torch.rand(2, 32, 32, 32, device=gpu)
    sb = cm.forward(xs, geom)
    ys = torch.rand(2, 24, 40, 48, device=gpu)
    ys[1] *= 3.0  # a different fixed-point unit for the adjoint of each stack
    vb = cm.backward(ys, geom)
    ab = cm.adjoint(ys, geom)
    for i in range(2):
        assert torch.equal(sb[i], cm.forward(xs[i].contiguous(), geom))
        assert torch.equal(vb[i], cm.backward(ys[i].contiguous(), geom))
        assert torch.equal(ab[i], cm.adjoint(ys[i].contiguous(), geom))


def test_reference_binding_signature(oracle, small, gpu):
    """curadon_ext.forward_3d / backward_3d with the reference's 15 positional arguments."""
    rec, geom = small
    vol = np.random.default_rng(8).random((32, 32, 32))
    sino = torch.zeros(24, 40, 48, device=gpu)
    cm.forward_3d(dev(vol, gpu), geom.vol_shape, geom.vol_spacing, geom.vol_offset, sino,
                  geom.angles[0], geom.angles[1], geom.angles[2], geom.det_shape,
                  geom.det_spacing, geom.det_offset, geom.det_rotation, geom.DSO, geom.DSD,
                  geom.COR)
    check_close(host(sino), oracle.fwd3d(rec, vol), "forward_3d binding")
    out = torch.zeros(32, 32, 32, device=gpu)
    cm.backward_3d(out, geom.vol_shape, geom.vol_spacing, geom.vol_offset, sino, geom.angles[0],
                   geom.angles[1], geom.angles[2], geom.det_shape, geom.det_spacing,
                   geom.det_offset, geom.det_rotation, geom.DSO, geom.DSD, geom.COR)
    check_close(host(out), oracle.bp3d(rec, host(sino)), "backward_3d binding")


@pytest.mark.parametrize("prec,tol", [(torch.float16, 2e-3), (torch.bfloat16, 1.5e-2)])
def test_half_storage_3d(oracle, small, gpu, prec, tol):
    rec, _ = small
    geom = as_cone_geometry(rec, vol_prec=prec, sino_prec=prec)
    vol = phantom.ellipsoids_3d(32, seed=1, dtype=np.float64)
    r = rel_l2(host(cm.forward(dev(vol, gpu, prec), geom)), oracle.fwd3d(rec, vol))
    y = np.random.default_rng(9).random((24, 40, 48))
    rb = rel_l2(host(cm.backward(dev(y, gpu, prec), geom)), oracle.bp3d(rec, y))
    ra = rel_l2(host(cm.adjoint(dev(y, gpu, prec), geom)), oracle.adj3d(rec, y))
    print(f"3d {prec}: fwd {r:.3e} bp {rb:.3e} adj {ra:.3e}")
    assert max(r, rb, ra) <= tol


@pytest.fixture(scope="module")
def big(gpu):
    rec = c3()
    geom = as_cone_geometry(rec)
    vol = phantom.ellipsoids_3d(256, seed=0)
    return rec, geom, vol


def test_c3_forward_subset(oracle, big, gpu):
    rec, geom, vol = big
    out = host(cm.forward(dev(vol, gpu), geom))
    got, ref = [], []
    for a0 in (0, 45, 90, 200):
        r = oracle.fwd3d(rec, vol.astype(np.float64), angles=(a0, a0 + 1))
        got.append(out[a0])
        ref.append(r[a0])
    check_close(np.stack(got), np.stack(ref), "c3 fwd3d (4 projections)")


def test_c3_backward_subset(oracle, big, gpu):
    rec, geom, vol = big
    y = np.random.default_rng(10).random((360, 384, 384)).astype(np.float32)
    yt = dev(y, gpu)
    for mode in ("reference", "matched"):
        out = host(cm.backward(yt, geom, mode=mode))
        got, ref = [], []
        for z0 in (0, 128, 255):
            r = oracle.bp3d(rec, y.astype(np.float64), slabs=(z0, z0 + 1),
                            matched=mode == "matched")
            got.append(out[z0])
            ref.append(r[z0])
        check_close(np.stack(got), np.stack(ref), f"c3 bp3d {mode} (3 slices)")


@pytest.mark.parametrize("a0,a1", [(0, 2), (89, 91), (200, 202)])
def test_c3_adjoint_blocks(oracle, big, gpu, a0, a1):
    """The matched adjoint at the 256^3 workload against the oracle's transposed march, on a
    sinogram that is zero outside two projections (the adjoint of that block)."""
    rec, geom, _ = big
    y = np.zeros((360, 384, 384), dtype=np.float32)
    y[a0:a1] = np.random.default_rng(30 + a0).random((a1 - a0, 384, 384))
    got = host(cm.adjoint(dev(y, gpu), geom))
    ref = oracle.adj3d(rec, y.astype(np.float64), angles=(a0, a1))
    check_close(got, ref, f"c3 adj3d (projections {a0}..{a1 - 1})")


def test_c3_full_adjoint_dot_and_linearity(big, gpu):
    """The adjoint over all 360 projections, where each tile's 32-bit sums are drained into
    64-bit registers several times: the dot test against the forward, and linearity over four
    blocks of 90 projections (each a call of its own, with its own fixed-point unit)."""
    rec, geom, _ = big
    g = torch.Generator(device="cpu").manual_seed(5)
    x = torch.rand(256, 256, 256, generator=g).to(gpu)
    y = (torch.rand(360, 384, 384, generator=g) - 0.3).to(gpu)
    lhs = torch.vdot(cm.forward(x, geom).double().ravel(), y.double().ravel())
    aty = cm.adjoint(y, geom)
    rhs = torch.vdot(x.double().ravel(), aty.double().ravel())
    r = abs(float(lhs - rhs)) / abs(float(lhs))
    print(f"c3 dot test: {r:.3e}")
    assert r <= 1e-5
    parts = torch.zeros_like(aty)
    for a0 in range(0, 360, 90):
        yb = torch.zeros_like(y)
        yb[a0:a0 + 90] = y[a0:a0 + 90]
        cm.adjoint(yb, parts, geom)
    r = float((parts.double() - aty.double()).norm() / aty.double().norm())
    print(f"c3 adjoint linearity: {r:.3e}")
    assert r <= 1e-6


@pytest.mark.parametrize("prec,tol", [(torch.float16, 1e-3), (torch.bfloat16, 6e-3)])
def test_c3_half_storage(oracle, big, gpu, prec, tol):
    """16-bit storage at the 256^3 workload: forward on four projections and the reference
    backprojection on three slices against the FP64 oracle of the unrounded inputs and of the
    stored ones."""
    rec, _, vol = big
    geom = as_cone_geometry(rec, vol_prec=prec, sino_prec=prec)
    x = dev(vol, gpu, prec)
    out = host(cm.forward(x, geom))
    xs = host(x)
    got, ref, refr = [], [], []
    for a0 in (0, 45, 90, 200):
        got.append(out[a0])
        ref.append(oracle.fwd3d(rec, vol.astype(np.float64), angles=(a0, a0 + 1))[a0])
        refr.append(oracle.fwd3d(rec, xs, angles=(a0, a0 + 1))[a0])
    r, rr = rel_l2(np.stack(got), np.stack(ref)), rel_l2(np.stack(got), np.stack(refr))
    print(f"c3 fwd3d {prec} (4 projections): rel-L2 {r:.3e} vs exact input, {rr:.3e} vs "
          f"stored input", flush=True)
    del out, x
    y = np.random.default_rng(31).random((360, 384, 384)).astype(np.float32)
    y16 = dev(y, gpu, prec)
    bv = host(cm.backward(y16, geom))
    ys = host(y16)
    gb, rb, rbr = [], [], []
    for z0 in (0, 128, 255):
        gb.append(bv[z0])
        rb.append(oracle.bp3d(rec, y.astype(np.float64), slabs=(z0, z0 + 1))[z0])
        rbr.append(oracle.bp3d(rec, ys, slabs=(z0, z0 + 1))[z0])
    b, br = rel_l2(np.stack(gb), np.stack(rb)), rel_l2(np.stack(gb), np.stack(rbr))
    print(f"c3 bp3d {prec} (3 slices): rel-L2 {b:.3e} vs exact input, {br:.3e} vs stored "
          f"input", flush=True)
    assert max(r, b) <= tol and max(rr, br) <= tol


def test_c3_dot_product(big, gpu):
    rec, geom, _ = big
    g = torch.Generator(device=gpu).manual_seed(11)
    x = torch.rand(256, 256, 256, device=gpu, generator=g)
    y = torch.rand(360, 384, 384, device=gpu, generator=g)
    lhs = torch.vdot(cm.forward(x, geom).double().ravel(), y.double().ravel())
    rhs = torch.vdot(x.double().ravel(), cm.adjoint(y, geom).double().ravel())
    r = abs(float(lhs - rhs)) / abs(float(lhs))
    print(f"c3 dot test: {r:.3e}")
    assert r <= 1e-5


def test_circular_orbit_odd_shapes(oracle, gpu):
    """Circular orbit (the backprojectors run along y): volume sizes that are no multiple of
    the run length or tile, anisotropic voxels, offsets, a panel turned about y."""
    rec = cone(30, 40, 34, 16, 70.0, 110.0, du=(1.1, 0.9), vol_spacing=(0.9, 1.2, 1.0),
               vol_offset=(0.7, -0.4, 0.3), det_offset=(1.2, -0.8), det_rotation=(0.03, 0.0, 0.0),
               cor=0.5, nz=19, ny=21)
    geom = as_cone_geometry(rec)
    y = np.random.default_rng(12).random((16, 34, 40))
    for mode in ("reference", "matched", "fdk"):
        ref = oracle.bp3d(rec, y, matched=mode == "matched", fdk=mode == "fdk")
        check_close(host(cm.backward(dev(y, gpu), geom, mode=mode)), ref,
                    f"circular odd-shape bp3d {mode}")
    yb = np.stack([y, 0.5 * y])
    out = host(cm.backward(dev(yb, gpu), geom, mode="fdk"))
    check_close(out[1], 0.5 * oracle.bp3d(rec, y, fdk=True), "circular odd-shape bp3d batch 1")


@pytest.mark.parametrize("dso,maxabs", [(34.0, 2e-5), (30.0, 5e-5)])
def test_circular_orbit_high_magnification(oracle, gpu, dso, maxabs):
    """Circular orbit with the source just outside the volume's bounding circle (radius 28.3
    for a 40-voxel-wide volume; DSD 60): tiles near the source have panel footprints far wider
    and taller than the staged LDS slot (64 x 60 cells) and are sampled from global memory,
    tiles far from it are staged; both against the oracle, in every mode. At DSO 30 the voxels
    next to the source sit 1.7 units from it: magnification about 35 and FDK weight
    (DSO / depth)^2 about 300 amplify the FP32 rounding of their panel coordinate, so the
    max-abs bound there is 5e-5 (rel-L2 stays below 1e-5)."""
    rec = cone(40, 96, 88, 12, dso, 60.0, ny=36)
    geom = as_cone_geometry(rec)
    y = np.random.default_rng(21).random((12, 88, 96))
    for mode in ("reference", "matched", "fdk"):
        ref = oracle.bp3d(rec, y, matched=mode == "matched", fdk=mode == "fdk")
        check_close(host(cm.backward(dev(y, gpu), geom, mode=mode)), ref,
                    f"high-magnification (DSO {dso}) bp3d {mode}", maxabs=maxabs)


@pytest.mark.parametrize("circular", [True, False])
def test_odd_panel_every_mode(oracle, gpu, circular):
    """A panel of 7 x 5 pixels and 3 projections: the staged panel stack holds
    (7 + 4) (5 + 4) 3 = 297 elements, not a multiple of four, so the staging kernel's last
    thread writes its tail element by element. Circular orbits take the y-run backprojector,
    tilted ones the column backprojector; every mode against the oracle."""
    if circular:
        rec = cone(9, 7, 5, 3, 30.0, 50.0, nz=7, ny=8)
    else:
        ang = np.array([[0.2, 1.9, 4.0], [0.1, -0.2, 0.15], [0.0, 0.1, -0.1]])
        rec = cone(9, 7, 5, 3, 30.0, 50.0, angles=ang, nz=7, ny=8)
    geom = as_cone_geometry(rec)
    y = np.random.default_rng(40).random((3, 5, 7))
    for mode in ("reference", "matched", "fdk"):
        ref = oracle.bp3d(rec, y, matched=mode == "matched", fdk=mode == "fdk")
        check_close(host(cm.backward(dev(y, gpu), geom, mode=mode)), ref,
                    f"odd panel ({'circular' if circular else 'tilted'}) bp3d {mode}")
