"""GPU tests of the continuous-batching FBP server (curadon_amd.serving) and of the native
multi-job backprojector behind it.

Job recipe (fan and cone): DSO = 2 n, DSD = 3 n (isocentre-to-detector n), A = 2 n angles over
2 pi; fan detector ceil(sqrt(2) n) cells, cone panel 1.5 n x 1.5 n. Raw projections are the
FP32 forward projections of seeded ellipse / ellipsoid phantoms.

Bars:
  * served volume == curadon_amd.fbp with the same storage precision, bit for bit;
  * a job's volume is bit-identical alone, in a mix, and split over many launches;
  * FP16 storage: rel-L2 <= 1e-3 against the FP64 pipeline (FP64 filtering + the oracle's
    weighted backprojector); BF16 storage: rel-L2 <= 5e-3;
  * admission never holds more than the HBM budget, and jobs that arrive while others run
    join their launches.
"""
import math

import numpy as np
import pytest
import torch

import curadon_amd as cm
from curadon_amd import phantom
from curadon_amd import reconstruction as F
from curadon_amd.serving import FBPJob, FBPServer, footprint
from geoms import cone, fan

pytestmark = pytest.mark.gpu


def fan_case(n, seed, storage=torch.float16):
    rec = fan(n, int(math.ceil(math.sqrt(2) * n)), 2 * n, 2.0 * n, 3.0 * n)
    g = cm.FanGeometry(DSD=rec.DSD, DSO=rec.DSO, angles=rec.angles, vol_shape=rec.vol_shape,
                       det_shape=rec.det_shape, vol_prec=storage, sino_prec=storage)
    g32 = cm.FanGeometry(DSD=rec.DSD, DSO=rec.DSO, angles=rec.angles, vol_shape=rec.vol_shape,
                         det_shape=rec.det_shape)
    img = torch.tensor(phantom.ellipses_2d(n, seed=seed), dtype=torch.float32, device="cuda")
    raw = cm.forward(img, g32)
    return rec, g, raw


def cone_case(n, seed, storage=torch.float16):
    p = int(1.5 * n)
    rec = cone(n, p, p, 2 * n, 2.0 * n, 3.0 * n)
    kw = dict(DSD=rec.DSD, DSO=rec.DSO, angles=rec.angles, vol_shape=rec.vol_shape,
              det_shape=rec.det_shape)
    g = cm.ConeGeometry(**kw, vol_prec=storage, sino_prec=storage)
    g32 = cm.ConeGeometry(**kw)
    vol = torch.tensor(phantom.ellipsoids_3d(n, seed=seed), device="cuda")
    raw = cm.forward(vol, g32)
    return rec, g, raw


def jobs_of(cases, arrivals=None):
    out = []
    for i, (_, g, raw) in enumerate(cases):
        out.append(FBPJob(geom=g, projections=raw, arrival=0.0 if arrivals is None
                          else arrivals[i], tag=f"job{i}"))
    torch.cuda.synchronize()
    return out


@pytest.fixture(scope="module")
def mix(gpu):
    return [fan_case(64, 1), cone_case(32, 2), fan_case(128, 3), cone_case(48, 4),
            fan_case(96, 5), fan_case(64, 6)]


def test_served_equals_plain_fbp(mix):
    srv = FBPServer()
    jobs = jobs_of(mix)
    st = srv.run(jobs)
    print(f"served {st['jobs']} jobs in {srv.rounds} launches, p50 {st['latency_ms_p50']:.2f} ms")
    for (rec, g, raw), j in zip(mix, jobs):
        ref = cm.fbp(raw, g)
        assert j.volume.dtype == torch.float16
        assert torch.equal(j.volume, ref), j.tag


def test_alone_equals_mixed_and_split(mix):
    target = 2  # the 128^2 fan job
    alone = jobs_of([mix[target]])
    FBPServer().run(alone)
    mixed = jobs_of(mix, arrivals=[0.0, 0.0, 0.001, 0.0, 0.002, 0.0005])
    FBPServer().run(mixed)
    # a round budget of one tile per job: every job is split over many launches
    split = jobs_of(mix)
    srv = FBPServer(round_cost=1.0, oldest_share=0.0)
    srv.run(split)
    assert srv.rounds >= 64
    assert torch.equal(alone[0].volume, mixed[target].volume)
    assert torch.equal(alone[0].volume, split[target].volume)
    for a, b in zip(mixed, split):
        assert torch.equal(a.volume, b.volume), a.tag


@pytest.mark.parametrize("storage,tol", [(torch.float16, 1e-3), (torch.bfloat16, 5e-3)])
def test_error_against_fp64_pipeline(oracle, gpu, storage, tol):
    cases = [fan_case(128, 7, storage), cone_case(48, 8, storage)]
    jobs = jobs_of(cases)
    FBPServer(storage=storage).run(jobs)
    for (rec, g, raw), j in zip(cases, jobs):
        q = F.fbp_filter(raw.double().cpu(), g).numpy()
        if isinstance(g, cm.FanGeometry):
            ref = oracle.bp2d(rec, q, wpow=2)
        else:
            ref = oracle.bp3d(rec, q, fdk=True)
        got = j.volume.double().cpu().numpy()
        r = float(np.linalg.norm(got - ref) / np.linalg.norm(ref))
        print(f"{type(g).__name__} {tuple(g.vol_shape)} {storage}: rel-L2 vs FP64 {r:.3e}")
        assert r <= tol


def test_admission_respects_budget(mix):
    jobs = jobs_of(mix)
    biggest = max(sum(footprint(j, torch.float16)) for j in jobs)
    srv = FBPServer(hbm_budget=biggest + 1024)
    srv.run(jobs)
    assert srv.max_used_seen <= srv.hbm_budget
    assert all(j.t_done >= j.t_admit for j in jobs)
    for (rec, g, raw), j in zip(mix, jobs):
        assert torch.equal(j.volume, cm.fbp(raw, g)), j.tag
    with pytest.raises(MemoryError):
        FBPServer(hbm_budget=1024).run(jobs_of(mix[:1]))


def test_late_arrivals_join_running_launches(gpu):
    big = cone_case(96, 9)
    small = fan_case(64, 10)
    jobs = jobs_of([big, small], arrivals=[0.0, 0.002])
    srv = FBPServer(round_cost=2048 * 192 * 2.0, oldest_share=0.0)
    srv.run(jobs)
    first_big = next(i for i, r in enumerate(srv.launch_log) if any(j == 0 for j, _, _ in r))
    shared = [i for i, r in enumerate(srv.launch_log)
              if {j for j, _, _ in r} == {0, 1}]
    print(f"{srv.rounds} launches; big job from launch {first_big}, both jobs in {len(shared)}")
    assert jobs[1].t_admit > 0.0
    assert shared and shared[0] > first_big
    assert jobs[1].t_done < jobs[0].t_done
    assert torch.equal(jobs[0].volume, cm.fbp(big[2], big[1]))
    assert torch.equal(jobs[1].volume, cm.fbp(small[2], small[1]))


def test_native_job_validation(gpu):
    from curadon_amd import _lib
    srv = _lib.Server(0, max_jobs=2)
    v = torch.zeros(8, 8, device="cuda", dtype=torch.float16)
    s = torch.zeros(4, 11, device="cuda", dtype=torch.float16)
    t = torch.zeros(4, 4, device="cuda")
    j = _lib.BPJob(kind=_lib.JOB_FAN2D, mode=_lib.BP2D_FBP, sino=s.data_ptr(),
                   table=t.data_ptr(), volume=v.data_ptr(), nx=8, ny=8, nz=1, nu=7, nv=1,
                   nangles=4, sx=1.0, sy=1.0, ox=0.0, oy=0.0, inv_du=1.0, tile_begin=0,
                   tile_end=2)
    assert _lib.bp_job_tiles(j) == 2
    srv.backproject([j], torch.float16, torch.float16)
    bad = _lib.BPJob.from_buffer_copy(j)
    bad.tile_end = 3
    with pytest.raises(ValueError):
        srv.backproject([bad], torch.float16, torch.float16)
    with pytest.raises(ValueError):
        srv.backproject([j, j, j], torch.float16, torch.float16)
    host = torch.zeros(8, 8, dtype=torch.float16)
    bad = _lib.BPJob.from_buffer_copy(j)
    bad.volume = host.data_ptr()
    with pytest.raises(AttributeError):
        srv.backproject([bad], torch.float16, torch.float16)
    torch.cuda.synchronize()


def _c5_geometry(kind, n, storage):
    A = 2 * n
    angles = np.linspace(0, 2 * np.pi, A, endpoint=False)
    if kind == "fan":
        return cm.FanGeometry(DSD=3.0 * n, DSO=2.0 * n, angles=angles, vol_shape=(n, n),
                              det_shape=int(math.ceil(math.sqrt(2) * n)), vol_prec=storage,
                              sino_prec=storage)
    p = (3 * n) // 2
    return cm.ConeGeometry(DSD=3.0 * n, DSO=2.0 * n, angles=angles, vol_shape=(n, n, n),
                           det_shape=(p, p), vol_prec=storage, sino_prec=storage)


def _fp64_subset(oracle, served, raw, g):
    """FP64 pipeline on a subset: two image rows (fan) or the central slab (cone), filtering
    only the panel rows that the slab projects onto (found from the FDK matrices: a
    linear-fractional map takes its extremes over the slab at the slab's corners)."""
    n = int(g.vol_shape[0])
    if isinstance(g, cm.FanGeometry):
        q = F.fbp_filter(raw.double().cpu(), g).numpy()
        rows = (n // 4, n // 2)
        ref = np.stack([oracle.bp2d(g, q, rows=(r, r + 1), wpow=2)[r] for r in rows])
        return served.double().cpu().numpy()[list(rows)], ref
    V, U = (int(v) for v in g.det_shape)
    z = n // 2
    M = g.plan.bp_matrices("fdk").astype(np.float64)
    c = (n - 1) / 2.0
    corners = np.array([[x, y, z - c, 1.0] for x in (-c, c) for y in (-c, c)])
    P = np.einsum("aij,kj->aki", M, corners)
    v = P[..., 1] / P[..., 2]
    v0, v1 = max(0, int(math.floor(v.min())) - 2), min(V, int(math.ceil(v.max())) + 3)
    H = F.ramp_response(U).double()
    cw = F.cosine_weights_cone(g).double()
    scale = (math.pi / g.nangles) * g.DSD / (float(g.det_spacing[1]) * g.DSO)
    q = np.zeros((g.nangles, V, U))
    q[:, v0:v1] = (F.filter_rows(raw[:, v0:v1].double().cpu() * cw[v0:v1], H) * scale).numpy()
    ref = oracle.bp3d(g, q, slabs=(z, z + 1), fdk=True)[z]
    return served[z].double().cpu().numpy(), ref


def test_every_job_class_of_the_mix_against_fp64(oracle, gpu):
    """One job of each class of the serving mix (fan n = 128..2048, cone n = 128..384, FP16
    storage), served together; rel-L2 against the FP64 pipeline on a subset <= 1e-3."""
    classes = [("fan", n) for n in (128, 256, 512, 1024, 2048)] + \
              [("cone", n) for n in (128, 256, 384)]
    jobs, raws = [], []
    for i, (kind, n) in enumerate(classes):
        g = _c5_geometry(kind, n, torch.float16)
        g32 = _c5_geometry(kind, n, 32)
        if kind == "fan":
            x = torch.tensor(phantom.ellipses_2d(n, seed=i), dtype=torch.float32, device="cuda")
        else:
            x = torch.tensor(phantom.ellipsoids_3d(n, seed=i), device="cuda")
        raws.append(cm.forward(x, g32))
        jobs.append(FBPJob(geom=g, projections=raws[-1], tag=f"{kind}{n}"))
    torch.cuda.synchronize()
    FBPServer(hbm_budget=64 << 30).run(jobs)
    worst = 0.0
    for j, raw in zip(jobs, raws):
        got, ref = _fp64_subset(oracle, j.volume, raw, j.geom)
        r = float(np.linalg.norm(got - ref) / np.linalg.norm(ref))
        print(f"{j.tag}: FP16 storage rel-L2 vs FP64 {r:.3e}", flush=True)
        worst = max(worst, r)
    assert worst <= 1e-3


@pytest.mark.parametrize("storage", [torch.float32, torch.float16, torch.bfloat16])
@pytest.mark.parametrize("offset", [0, 1, 3])
def test_stage_projections_tail_and_unaligned_destination(gpu, storage, offset):
    """Staging into a destination that starts `offset` elements into its buffer (no vector
    stores there) and whose element count, A (D + 4) = 3 * 9 = 27 (fan) and
    A (V + 4) (U + 4) = 3 * 7 * 9 = 189 (cone), is no multiple of four (the last thread's tail):
    every element, borders included, equals the host-side expectation."""
    from curadon_amd import _lib
    g = torch.Generator().manual_seed(50 + offset)
    for shape in ((3, 5), (3, 3, 5)):
        src = torch.rand(shape, generator=g).to(gpu)
        rs = torch.rand(shape[0], generator=g).to(gpu)
        staged_shape = (3, 9) if len(shape) == 2 else (3, 7, 9)
        n = int(np.prod(staged_shape))
        buf = torch.full((n + offset + 4,), 7.0, dtype=storage, device=gpu)
        dst = buf[offset:offset + n].view(staged_shape)
        keep = _lib.stage_projections(src, dst, row_scale=rs, scale=0.5)
        torch.cuda.synchronize()
        del keep
        # FP32 products as the kernel forms them, then one rounding to the storage type
        exp = torch.zeros(staged_shape, dtype=torch.float32)
        val = (src.cpu() * 0.5) * rs.cpu().reshape((-1,) + (1,) * (len(shape) - 1))
        if len(shape) == 2:
            exp[:, 2:7] = val
        else:
            exp[:, 2:5, 2:7] = val
        exp = exp.to(storage)
        assert torch.equal(dst.cpu(), exp), (shape, offset)
        # nothing written before or after the destination
        assert torch.all(buf[:offset].cpu() == 7.0) and torch.all(buf[offset + n:].cpu() == 7.0)
