"""Differentiable batched 2D fan beam: 16 images of 512^2 FP32, 724 cells, 512 angles over
2 pi, DSO 1024, DSD 1536. loss = ||A x - y||^2 with FP64 reduction, y = A(x + 0.05 p) + N(0,
0.01^2) where p is a second ellipse phantom. The gradient comes from TorchForward's backward
(the hand-written adjoint) and is checked against
  * central finite differences, h = 1e-3, at 64 random pixels: rel-L2 over the 64 entries
    <= 1e-2 (the FP32 forward's rounding in the perturbed rays bounds the differences at
    ~1e-3 of the loss change);
  * the FP64 gradient 2 A^T (A x - y) of the oracle for one image of the batch: rel-L2 <= 1e-5.
"""
import numpy as np
import pytest
import torch

import curadon_amd as cm
from curadon_amd import phantom
from geoms import as_fan_geometry, c4

pytestmark = pytest.mark.gpu

B = 16


@pytest.fixture(scope="module")
def problem(gpu):
    rec = c4()
    geom = as_fan_geometry(rec)
    x = np.stack([phantom.ellipses_2d(512, seed=s) for s in range(B)])
    p = np.stack([phantom.ellipses_2d(512, seed=100 + s) for s in range(B)])
    xt = torch.tensor(x, dtype=torch.float32, device=gpu)
    xp = torch.tensor(x + 0.05 * p, dtype=torch.float32, device=gpu)
    g = torch.Generator(device=gpu).manual_seed(3)
    y = cm.forward(xp, geom) + 0.01 * torch.randn(B, 512, 724, device=gpu, generator=g)
    return rec, geom, xt, y


def loss_fn(geom, x, y):
    r = cm.TorchForward.apply(x, geom) - y
    return (r.double() ** 2).sum()


def test_batched_gradient_finite_differences(problem, gpu):
    rec, geom, x, y = problem
    xg = x.clone().requires_grad_(True)
    loss = loss_fn(geom, xg, y)
    loss.backward()
    grad = xg.grad.double()
    assert grad.shape == x.shape
    r = np.random.default_rng(4)
    idx = [(int(r.integers(B)), int(r.integers(64, 448)), int(r.integers(64, 448)))
           for _ in range(64)]
    h = 1e-3
    fd, ag = [], []
    with torch.no_grad():
        for b, i, j in idx:
            xb = x[b:b + 1].clone()
            xb[0, i, j] += h
            lp = loss_fn(geom, xb, y[b:b + 1])
            xb[0, i, j] -= 2 * h
            lm = loss_fn(geom, xb, y[b:b + 1])
            fd.append(float(lp - lm) / (2 * h))
            ag.append(float(grad[b, i, j]))
    fd, ag = np.array(fd), np.array(ag)
    err = np.linalg.norm(fd - ag) / np.linalg.norm(ag)
    print(f"c4 finite differences vs autograd (64 pixels): rel-L2 {err:.3e}")
    assert err <= 1e-2


def test_batched_gradient_matches_fp64(oracle, problem, gpu):
    rec, geom, x, y = problem
    xg = x.clone().requires_grad_(True)
    loss_fn(geom, xg, y).backward()
    b = 5
    resid = (cm.forward(x[b].contiguous(), geom) - y[b]).double().cpu().numpy()
    ref_adj = 2.0 * oracle.adj2d(rec, resid)
    got = xg.grad[b].double().cpu().numpy()
    err = np.linalg.norm(got - ref_adj) / np.linalg.norm(ref_adj)
    print(f"c4 gradient vs FP64 adjoint of the same residual: rel-L2 {err:.3e}")
    assert err <= 1e-5
    # the full FP64 gradient, residual included
    ref_full = 2.0 * oracle.adj2d(
        rec, oracle.fwd2d(rec, x[b].double().cpu().numpy()) - y[b].double().cpu().numpy())
    err_full = np.linalg.norm(got - ref_full) / np.linalg.norm(ref_full)
    print(f"c4 gradient vs FP64 gradient 2 A^T (A x - y): rel-L2 {err_full:.3e}")
    assert err_full <= 1e-4


def test_gradients_of_all_autograd_functions(gpu):
    """TorchForward / TorchAdjoint form an exact pair; TorchBackprojection's gradient is the
    transpose of the backprojection; each returns one gradient per input."""
    from geoms import fan
    geom = as_fan_geometry(fan(48, 70, 40, 96.0, 144.0))
    x = torch.rand(48, 48, device=gpu, requires_grad=True)
    s = torch.rand(40, 70, device=gpu, requires_grad=True)
    ws = torch.rand(40, 70, device=gpu)
    wx = torch.rand(48, 48, device=gpu)
    (cm.TorchForward.apply(x, geom) * ws).sum().backward()
    torch.testing.assert_close(x.grad, cm.adjoint(ws, geom), rtol=1e-6, atol=1e-6)
    (cm.TorchAdjoint.apply(s, geom) * wx).sum().backward()
    torch.testing.assert_close(s.grad, cm.forward(wx, geom), rtol=1e-6, atol=1e-6)
    s.grad = None
    (cm.TorchBackprojection.apply(s, geom) * wx).sum().backward()
    assert torch.equal(s.grad, cm.backward_transpose(wx, geom))
