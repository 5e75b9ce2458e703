"""GPU parity of the 2D fan-beam HIP kernels against the FP64 oracle, through the C-ABI.

Tolerances (FP32 storage and arithmetic vs FP64):
  forward   rel-L2 <= 1e-5, max-abs <= 2e-5 * max|ref|
  backward  rel-L2 <= 1e-5, max-abs <= 2e-5 * max|ref|
  adjoint   rel-L2 <= 1e-5, max-abs <= 2e-5 * max|ref|
  dot test  |<Ax,y> - <x,A^T y>| / |<Ax,y>| <= 1e-5 (both sides on the GPU in FP32)
At the 1024^2 workload the oracle runs on row / angle subsets (the operators are per-ray and
per-pixel, so a subset is an exact restriction).
"""
import numpy as np
import pytest
import torch

import curadon_amd as cm
from curadon_amd import phantom
from geoms import as_fan_geometry, c1, c2, c4, fan, tiny

pytestmark = pytest.mark.gpu

RTOL_L2 = 1e-5
MAXABS = 2e-5


def rel_l2(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def check_close(got, ref, what, rtol=RTOL_L2, maxabs=MAXABS):
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
def g1(gpu):
    rec = c1()
    return rec, as_fan_geometry(rec)


def test_c1_forward(oracle, g1, gpu):
    rec, geom = g1
    img = phantom.ellipses_2d(64, seed=0)
    ref = oracle.fwd2d(rec, img)
    out = cm.forward(dev(img, gpu), geom)
    torch.cuda.synchronize()
    check_close(host(out), ref, "c1 forward")


def test_c1_backward(oracle, g1, gpu):
    rec, geom = g1
    sino = np.random.default_rng(1).random((360, 91))
    ref = oracle.bp2d(rec, sino)
    vol = torch.zeros(64, 64, device=gpu)
    cm.backward(dev(sino, gpu), vol, geom)
    check_close(host(vol), ref, "c1 backward")


def test_c1_backward_accumulates(g1, gpu):
    rec, geom = g1
    s = dev(np.random.default_rng(2).random((360, 91)), gpu)
    base = torch.rand(64, 64, device=gpu)
    once = cm.backward(s, geom)
    acc = cm.backward(s, base.clone(), geom)
    torch.testing.assert_close(acc, base + once, rtol=1e-6, atol=1e-5)


def test_c1_adjoint(oracle, g1, gpu):
    rec, geom = g1
    sino = np.random.default_rng(3).random((360, 91))
    ref = oracle.adj2d(rec, sino)
    out = cm.adjoint(dev(sino, gpu), geom)
    check_close(host(out), ref, "c1 adjoint")


def test_c1_dot_product(g1, gpu):
    rec, geom = g1
    g = torch.Generator(device="cpu").manual_seed(4)
    x = torch.rand(64, 64, generator=g).to(gpu)
    y = torch.rand(360, 91, generator=g).to(gpu)
    lhs = torch.vdot(cm.forward(x, geom).double().ravel(), y.double().ravel())
    rhs = torch.vdot(x.double().ravel(), cm.adjoint(y, geom).double().ravel())
    r = abs(float(lhs - rhs)) / abs(float(lhs))
    print(f"c1 dot test: {r:.3e}")
    assert r <= 1e-5


def test_tiny_golden(g1, gpu):
    import os
    d = np.load(os.path.join(os.path.dirname(__file__), "golden", "fan2d_tiny.npz"))
    geom = as_fan_geometry(tiny())
    check_close(host(cm.forward(dev(d["image"], gpu), geom)), d["sino"], "tiny forward")
    check_close(host(cm.backward(dev(d["sino_in"], gpu), geom)), d["bp"], "tiny backward")
    check_close(host(cm.adjoint(dev(d["sino_in"], gpu), geom)), d["adj"], "tiny adjoint")


@pytest.fixture(scope="module")
def g2(gpu):
    rec = c2()
    geom = as_fan_geometry(rec)
    img = phantom.ellipses_2d(1024, seed=0)
    return rec, geom, img


ANGLE_BLOCKS = [(0, 6), (87, 93), (178, 184), (357, 363), (534, 540), (714, 720)]
ROW_BLOCKS = [(0, 4), (300, 304), (510, 514), (1020, 1024)]


def test_c2_forward_subsets(oracle, g2, gpu):
    rec, geom, img = g2
    out = host(cm.forward(dev(img, gpu), geom))
    got, ref = [], []
    for a0, a1 in ANGLE_BLOCKS:
        r = oracle.fwd2d(rec, img, angles=(a0, a1))
        got.append(out[a0:a1])
        ref.append(r[a0:a1])
    check_close(np.concatenate(got), np.concatenate(ref), "c2 forward (36 angles)")


def test_c2_backward_subsets(oracle, g2, gpu):
    rec, geom, img = g2
    sino = host(cm.forward(dev(img, gpu), geom))
    vol = host(cm.backward(dev(sino, gpu), geom))
    got, ref = [], []
    for y0, y1 in ROW_BLOCKS:
        r = oracle.bp2d(rec, sino, rows=(y0, y1))
        got.append(vol[y0:y1])
        ref.append(r[y0:y1])
    check_close(np.concatenate(got), np.concatenate(ref), "c2 backward (16 rows)")


# Blocks of six angles where the fan's rays change dominant axis (the fan's half angle is
# 13.3 degrees, so each block holds rays of both kinds near 45, 135, 225 and 315 degrees) and
# where every ray has the same dominant axis (0, 90 degrees).
ADJ_BLOCKS = [(0, 6), (87, 93), (177, 183), (264, 270), (447, 453), (627, 633)]


@pytest.mark.parametrize("a0,a1", ADJ_BLOCKS)
def test_c2_adjoint_angle_blocks(oracle, g2, gpu, a0, a1):
    """Adjoint of six-angle sinogram blocks at 1024^2: rel-L2 <= 1e-5 and max-abs <= 5e-5 of
    max|ref|. A texel's value here is a sum of about a dozen samples, each weighted at an FP32
    position (the other coordinate is fma(j, v, m) with |j| up to 700: about 3e-5 texels of
    rounding), so the texel-wise bound is looser than for the full-angle operators, whose sums
    average that rounding out."""
    rec, geom, _ = g2
    y = np.zeros((720, 1449))
    y[a0:a1] = np.random.default_rng(5 + a0).random((a1 - a0, 1449))
    ref = oracle.adj2d(rec, y, angles=(a0, a1))
    got = host(cm.adjoint(dev(y, gpu), geom))
    check_close(got, ref, f"c2 adjoint (angles {a0}..{a1 - 1})", maxabs=5e-5)


def test_c2_dot_product(g2, gpu):
    rec, geom, img = g2
    g = torch.Generator(device="cpu").manual_seed(6)
    x = torch.rand(1024, 1024, generator=g).to(gpu)
    y = torch.rand(720, 1449, generator=g).to(gpu)
    lhs = torch.vdot(cm.forward(x, geom).double().ravel(), y.double().ravel())
    rhs = torch.vdot(x.double().ravel(), cm.adjoint(y, geom).double().ravel())
    r = abs(float(lhs - rhs)) / abs(float(lhs))
    print(f"c2 dot test: {r:.3e}")
    assert r <= 1e-5


def test_c2_adjoint_linearity_over_blocks(g2, gpu):
    """The full adjoint (each workgroup's 32-bit tile sums drained into 64-bit registers several
    times) equals the sum of the adjoints of four blocks of 180 angles, each a call of its own
    with its own fixed-point unit, to the rounding of the terms."""
    rec, geom, img = g2
    g = torch.Generator(device="cpu").manual_seed(8)
    y = (torch.rand(720, 1449, generator=g) - 0.3).to(gpu)
    full = cm.adjoint(y, geom).double()
    parts = torch.zeros(1024, 1024, device=gpu)
    for a0 in range(0, 720, 180):
        yb = torch.zeros_like(y)
        yb[a0:a0 + 180] = y[a0:a0 + 180]
        cm.adjoint(yb, parts, geom)
    r = float((parts.double() - full).norm() / full.norm())
    print(f"c2 adjoint linearity: {r:.3e}")
    assert r <= 1e-6


def test_total_samples_match_oracle(oracle, g2):
    rec, geom, img = g2
    _, n = oracle.fwd2d(rec, np.zeros((1024, 1024)), angles=(0, 720), count=True)
    assert geom.plan.info()["total_samples"] == n


def test_forward_deterministic_and_block_independent(g1, gpu):
    rec, geom = g1
    x = torch.rand(64, 64, device=gpu)
    a = cm.forward(x, geom)
    b = cm.forward(x, geom)
    plan = geom.plan
    old = (plan.forward_block_x, plan.forward_block_y)
    try:
        plan.set_block(128, 2)
        c = cm.forward(x, geom)
    finally:
        plan.set_block(*old)
    assert torch.equal(a, b) and torch.equal(a, c)
    y = torch.rand(360, 91, device=gpu)
    assert torch.equal(cm.adjoint(y, geom), cm.adjoint(y, geom))


def test_forward_wide_bands_all_block_shapes(oracle, gpu):
    """Detector cells 2.5 texels apart: a 64-ray workgroup's bands are about 70 to 160 columns
    wide, so some are staged in LDS and some sampled from global memory, and a 128-ray
    workgroup's wide bands all take the global path. Every launch shape gives the same sums bit
    for bit, and they match the oracle."""
    rec = fan(256, 160, 90, 512.0, 768.0, du=2.5)
    geom = as_fan_geometry(rec)
    img = phantom.ellipses_2d(256, seed=4)
    x = dev(img, gpu)
    ref = oracle.fwd2d(rec, img)
    plan = geom.plan
    old = (plan.forward_block_x, plan.forward_block_y)
    outs = []
    try:
        for bx, by in ((64, 1), (128, 1), (64, 2)):
            plan.set_block(bx, by)
            outs.append(cm.forward(x, geom))
    finally:
        plan.set_block(*old)
    for o in outs[1:]:
        assert torch.equal(outs[0], o)
    check_close(host(outs[0]), ref, "wide-band forward")


def test_block_shape_limit(g1):
    """Workgroups above the kernels' 256-thread launch bound are refused on the host."""
    plan = g1[1].plan
    old = (plan.forward_block_x, plan.forward_block_y)
    with pytest.raises(ValueError):
        plan.forward_block_y = 8          # 64 x 8 = 512 threads
    with pytest.raises(ValueError):
        plan.forward_block_x = 96         # not a multiple of 64
    assert (plan.forward_block_x, plan.forward_block_y) == old


def test_batched_equals_single(gpu):
    rec = fan(96, 136, 64, 192.0, 288.0)
    geom = as_fan_geometry(rec)
    xs = torch.rand(3, 96, 96, device=gpu)
    sb = cm.forward(xs, geom)
    for i in range(3):
        assert torch.equal(sb[i], cm.forward(xs[i].contiguous(), geom))
    ys = torch.rand(3, 64, 136, device=gpu)
    ys[2] *= 5.0  # a different fixed-point unit for the adjoint of each sinogram
    vb = cm.backward(ys, geom)
    ab = cm.adjoint(ys, geom)
    for i in range(3):
        assert torch.equal(vb[i], cm.backward(ys[i].contiguous(), geom))
        assert torch.equal(ab[i], cm.adjoint(ys[i].contiguous(), geom))


@pytest.mark.parametrize("prec,tol", [(torch.float16, 2e-3), (torch.bfloat16, 1.5e-2)])
def test_half_storage_c1(oracle, gpu, prec, tol):
    rec = c1()
    geom = as_fan_geometry(rec, vol_prec=prec, sino_prec=prec)
    img = phantom.ellipses_2d(64, seed=0)
    ref = oracle.fwd2d(rec, img)
    out = host(cm.forward(dev(img, gpu, prec), geom))
    r = rel_l2(out, ref)
    print(f"c1 forward {prec}: rel-L2 {r:.3e}")
    assert r <= tol
    sino = np.random.default_rng(7).random((360, 91))
    refb = oracle.bp2d(rec, sino)
    outb = host(cm.backward(dev(sino, gpu, prec), geom))
    rb = rel_l2(outb, refb)
    print(f"c1 backward {prec}: rel-L2 {rb:.3e}")
    assert rb <= tol


# 16-bit storage at the 1024^2 workload: the error against the FP64 oracle of the unrounded
# inputs (what a caller gets) and against the oracle of the rounded inputs (the operator's own
# part: FP32 arithmetic and the rounding of its output).
@pytest.mark.parametrize("prec,tol", [(torch.float16, 1e-3), (torch.bfloat16, 6e-3)])
def test_half_storage_c2(oracle, g2, gpu, prec, tol):
    rec, _, img = g2
    geom = as_fan_geometry(rec, vol_prec=prec, sino_prec=prec)
    x = dev(img, gpu, prec)
    xr = host(x)  # the image as stored
    out = host(cm.forward(x, geom))
    got, ref, refr = [], [], []
    for a0, a1 in ANGLE_BLOCKS[:4]:
        got.append(out[a0:a1])
        ref.append(oracle.fwd2d(rec, img, angles=(a0, a1))[a0:a1])
        refr.append(oracle.fwd2d(rec, xr, angles=(a0, a1))[a0:a1])
    got, ref, refr = np.concatenate(got), np.concatenate(ref), np.concatenate(refr)
    r, rr = rel_l2(got, ref), rel_l2(got, refr)
    print(f"c2 forward {prec} (24 angles): rel-L2 {r:.3e} vs exact input, {rr:.3e} vs "
          f"stored input", flush=True)
    sino = np.random.default_rng(8).random((720, 1449))
    s16 = dev(sino, gpu, prec)
    vol = host(cm.backward(s16, geom))
    gb, rb, rbr = [], [], []
    for y0, y1 in ROW_BLOCKS:
        gb.append(vol[y0:y1])
        rb.append(oracle.bp2d(rec, sino, rows=(y0, y1))[y0:y1])
        rbr.append(oracle.bp2d(rec, host(s16), rows=(y0, y1))[y0:y1])
    gb, rb, rbr = np.concatenate(gb), np.concatenate(rb), np.concatenate(rbr)
    b, br = rel_l2(gb, rb), rel_l2(gb, rbr)
    print(f"c2 backward {prec} (16 rows): rel-L2 {b:.3e} vs exact input, {br:.3e} vs "
          f"stored input", flush=True)
    assert max(r, b) <= tol
    assert max(rr, br) <= tol


def test_errors_like_reference(g1, gpu):
    rec, geom = g1
    x = torch.rand(64, 64, device=gpu)
    with pytest.raises(AttributeError, match="precision"):
        cm.forward(x.half(), geom)
    with pytest.raises(TypeError):
        cm.forward(x.double(), geom)
    with pytest.raises(ValueError):
        cm.forward(torch.rand(63, 64, device=gpu), geom)
    with pytest.raises(TypeError):
        cm.forward(x.t(), geom)  # not contiguous
    with pytest.raises(ValueError, match="Unknown geometry"):
        cm.forward(x, torch.zeros(1), None)


def _fan_case(rec, gpu, seed):
    geom = as_fan_geometry(rec)
    ny, nx = int(rec.vol_shape[0]), int(rec.vol_shape[1])
    r = np.random.default_rng(seed)
    x = r.random((ny, nx))
    y = r.random((rec.nangles, rec.det_shape))
    return geom, x, y


@pytest.mark.parametrize("case", ["anisotropic", "wide_cells", "degenerate", "short_arc"])
def test_unusual_geometries(oracle, gpu, case):
    """Forward, backprojection and adjoint against the oracle on geometries the workloads do
    not reach: a non-square odd image with anisotropic spacing, offsets, a detector offset and
    per-angle DSO / DSD; cells 3 texels wide on a 256-wide image, so that the 128 rays of a
    workgroup span about 230 columns and the forward samples those bands from global memory
    (the LDS path holds at most 188); a 2 x 3 image seen by one cell at one angle; and a short
    arc whose rays are all x-dominant."""
    if case == "anisotropic":
        rec = fan(45, 77, 50, 90.0, 150.0, du=1.3, det_offset=1.5, spacing=(0.7, 1.3),
                  offset=(2.0, -1.0), ny=31)
        rec.DSO = np.linspace(85.0, 95.0, 50)
        rec.DSD = rec.DSO + np.linspace(55.0, 65.0, 50)
    elif case == "wide_cells":
        rec = fan(256, 140, 16, 400.0, 667.0, du=3.0)
    elif case == "degenerate":
        rec = fan(3, 1, 1, 20.0, 35.0, ny=2, angles=[0.3])
    else:
        rec = fan(64, 91, 24, 128.0, 192.0, angles=np.linspace(-0.3, 0.3, 24) + np.pi / 2)
    geom, x, y = _fan_case(rec, gpu, 31)
    check_close(host(cm.forward(dev(x, gpu), geom)), oracle.fwd2d(rec, x), f"{case} fwd2d")
    check_close(host(cm.backward(dev(y, gpu), geom)), oracle.bp2d(rec, y), f"{case} bp2d")
    check_close(host(cm.adjoint(dev(y, gpu), geom)), oracle.adj2d(rec, y), f"{case} adj2d")
