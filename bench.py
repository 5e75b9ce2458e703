#!/usr/bin/env python3
"""Benchmark of the 2D fan-beam projection pair on one MI355X.

Workload (one "step" = one forward projection + one voxel-driven backprojection):
  image 1024 x 1024 FP32 (10 random ellipses, seed 0), pixel spacing 1; flat detector of 1449
  cells, spacing 1; 720 angles over [0, 2 pi); source-to-isocentre 2048, source-to-detector
  3072. Inputs are resident in HBM before timing starts.

Prints one JSON line: value = forward+backward pairs per second (wall clock over K steps,
synchronised on both sides), per-op device times from HIP events on the stream the kernels
run on, the rooflines of the dominant kernel (VALU issue, LDS and HBM, from committed per-launch
counter counts over the live op time), and a CPU baseline that times the oracle's FP32 port of
the reference arithmetic on the same workload.

Usage: python bench.py [--gpus 1] [--steps 50] [--warmup 3]
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

# MI355X peaks: 256 CUs at 2.4 GHz; FP32 VALU 128 lanes per clock per CU (157.3 TFLOP/s as FMA),
# LDS 128 bytes per clock per CU, HBM3E 8 TB/s
CUS, CLOCK_HZ = 256, 2.4e9
VALU_PEAK_LANES = 128 * CUS * CLOCK_HZ
LDS_PEAK_BPS = 128 * CUS * CLOCK_HZ
HBM_PEAK_BPS = 8.0e12
COUNTERS_FILE = os.path.join(ROOT, "profiles", "counters_fan2d_1024.json")

WORKLOAD = dict(n=1024, D=1449, A=720, DSO=2048.0, DSD=3072.0)


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def build_case(torch, cm, phantom, np):
    w = WORKLOAD
    angles = np.linspace(0, 2 * np.pi, w["A"], endpoint=False)
    geom = cm.FanGeometry(DSD=w["DSD"], DSO=w["DSO"], angles=angles,
                          vol_shape=(w["n"], w["n"]), det_shape=w["D"])
    img = torch.tensor(phantom.ellipses_2d(w["n"], seed=0), dtype=torch.float32,
                       device="cuda:0")
    sino = torch.empty(w["A"], w["D"], dtype=torch.float32, device="cuda:0")
    vol = torch.zeros(w["n"], w["n"], dtype=torch.float32, device="cuda:0")
    return geom, img, sino, vol


def tap_bytes(info, w, s=4):
    """Bytes a texture-less gather moves if every interpolation tap came from HBM:
    forward  sum_rays N * 4 taps * s + A * D * s
    backward A * n^2 * 2 taps * s + 2 * n^2 * s (read-modify-write of the image)."""
    n2 = w["n"] * w["n"]
    fwd = info["total_samples"] * 4 * s + w["A"] * w["D"] * s
    bwd = w["A"] * n2 * 2 * s + 2 * n2 * s
    return fwd, bwd


def cold_calls(torch, cm, phantom, np):
    """Plan construction and first calls on fresh geometries, wall clock with synchronisation:
    the 1024^2 fan beam (plan with its FP64 per-ray table, first forward, first
    backprojection) and the 256^3 cone beam (plan with its per-ray sample count, first forward).
    The warm figures elsewhere exclude this one-off plan work."""
    w = WORKLOAD
    out = {}
    img = torch.tensor(phantom.ellipses_2d(w["n"], seed=1), dtype=torch.float32, device="cuda:0")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    g = cm.FanGeometry(DSD=w["DSD"] + 1.0, DSO=w["DSO"],
                       angles=np.linspace(0, 2 * np.pi, w["A"], endpoint=False),
                       vol_shape=(w["n"], w["n"]), det_shape=w["D"])
    g.plan
    t1 = time.perf_counter()
    sino = cm.forward(img, g)
    torch.cuda.synchronize()
    t2 = time.perf_counter()
    cm.backward(sino, g)
    torch.cuda.synchronize()
    t3 = time.perf_counter()
    out["fan2d_1024"] = {"plan_ms": (t1 - t0) * 1e3, "first_fwd_ms": (t2 - t1) * 1e3,
                         "first_bwd_ms": (t3 - t2) * 1e3, "plan_and_first_pair_ms": (t3 - t0) * 1e3}
    vol = torch.tensor(phantom.ellipsoids_3d(256, seed=1), device="cuda:0")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    gc = cm.ConeGeometry(DSD=769.0, DSO=512.0, angles=np.linspace(0, 2 * np.pi, 360, endpoint=False),
                         vol_shape=(256, 256, 256), det_shape=(384, 384))
    gc.plan
    t1 = time.perf_counter()
    cm.forward(vol, gc)
    torch.cuda.synchronize()
    t2 = time.perf_counter()
    out["cone3d_256"] = {"plan_ms": (t1 - t0) * 1e3, "first_fwd_ms": (t2 - t1) * 1e3}
    return out


def time_ops(torch, stream, fns, reps):
    """Mean ms per call of each callable, HIP events on the launch stream."""
    out = {}
    for name, fn in fns:
        fn()
        torch.cuda.synchronize()
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record(stream)
        for _ in range(reps):
            fn()
        b.record(stream)
        torch.cuda.synchronize()
        out[name] = a.elapsed_time(b) / reps
    return out


def cone_case(torch, cm, phantom, np, stream, reps):
    """256^3 FP32 volume of 20 ellipsoids, 384 x 384 panel, 360 angles, DSO 512, DSD 768."""
    angles = np.linspace(0, 2 * np.pi, 360, endpoint=False)
    geom = cm.ConeGeometry(DSD=768.0, DSO=512.0, angles=angles, vol_shape=(256, 256, 256),
                           det_shape=(384, 384))
    vol = torch.tensor(phantom.ellipsoids_3d(256, seed=0), device="cuda:0")
    sino = torch.empty(360, 384, 384, device="cuda:0")
    bvol = torch.zeros_like(vol)
    info = geom.plan.info()
    t = time_ops(torch, stream, [
        ("fwd3d", lambda: cm.forward(vol, sino, geom)),
        ("bp3d", lambda: cm.backward(sino, bvol, geom)),
        ("bp3d_matched", lambda: cm.backward(sino, bvol, geom, mode="matched")),
        ("adj3d", lambda: cm.adjoint(sino, bvol, geom)),
        ("bp3d_transpose", lambda: cm.backward_transpose(vol, geom, sino)),
    ], reps)
    n3 = 256 ** 3
    fwd_b = info["total_samples"] * 8 * 4 + 360 * 384 * 384 * 4
    bp_b = 360 * n3 * 4 * 4 + 2 * n3 * 4
    return {
        "workload": "cone3d 256^3 f32, 384x384 panel, 360 angles, DSO 512, DSD 768",
        "total_samples": info["total_samples"],
        "fwd3d": {"ms": t["fwd3d"], "tap_equiv_GB_per_s": fwd_b / t["fwd3d"] / 1e6},
        "bp3d": {"ms": t["bp3d"], "tap_equiv_GB_per_s": bp_b / t["bp3d"] / 1e6},
        "bp3d_matched": {"ms": t["bp3d_matched"], "tap_equiv_GB_per_s": bp_b / t["bp3d_matched"] / 1e6},
        "adj3d": {"ms": t["adj3d"]},
        "bp3d_transpose": {"ms": t["bp3d_transpose"]},
        "fwd_bwd_pairs_per_s": 1e3 / (t["fwd3d"] + t["bp3d"]),
    }


def batch_case(torch, cm, phantom, np, stream, reps):
    """16 images of 512^2 FP32, 724 cells, 512 angles, DSO 1024, DSD 1536, one batched call."""
    angles = np.linspace(0, 2 * np.pi, 512, endpoint=False)
    geom = cm.FanGeometry(DSD=1536.0, DSO=1024.0, angles=angles, vol_shape=(512, 512),
                          det_shape=724)
    x = torch.stack([torch.tensor(phantom.ellipses_2d(512, seed=s), dtype=torch.float32)
                     for s in range(16)]).to("cuda:0")
    sino = torch.empty(16, 512, 724, device="cuda:0")
    g = torch.zeros_like(x)
    t = time_ops(torch, stream, [
        ("fwd", lambda: cm.forward(x, sino, geom)),
        ("adj", lambda: cm.adjoint(sino, g, geom)),
        ("bwd", lambda: cm.backward(sino, g, geom)),
    ], reps)
    return {"workload": "fan2d batch 16 x 512^2 f32, 724 cells, 512 angles (one call)",
            "fwd_ms": t["fwd"], "adj_ms": t["adj"], "bwd_ms": t["bwd"],
            "fwd_adj_images_per_s": 16e3 / (t["fwd"] + t["adj"])}


# ----------------------------------------------------------------------------- serving ---
# FBP / FDK job mix: 70 % fan-beam jobs, n in {128, 256, 512, 1024, 2048}, D = ceil(sqrt(2) n);
# 30 % cone-beam jobs, n in {128, 256, 384}, panel 1.5 n x 1.5 n; A = 2 n angles over 2 pi;
# DSO = 2 n, DSD = 3 n. Each job: cosine weights, ramp filter, weighted backprojection.
SERVING_FAN = (128, 256, 512, 1024, 2048)
SERVING_CONE = (128, 256, 384)


def serving_geometry(cm, np, kind, n, storage):
    import math
    A = 2 * n
    angles = np.linspace(0, 2 * np.pi, A, endpoint=False)
    if kind == "fan":
        return cm.FanGeometry(DSD=3.0 * n, DSO=2.0 * n, angles=angles, vol_shape=(n, n),
                              det_shape=int(math.ceil(math.sqrt(2) * n)), vol_prec=storage,
                              sino_prec=storage)
    p = (3 * n) // 2
    return cm.ConeGeometry(DSD=3.0 * n, DSO=2.0 * n, angles=angles, vol_shape=(n, n, n),
                           det_shape=(p, p), vol_prec=storage, sino_prec=storage)


def serving_case(torch, cm, phantom, np, n_jobs, seed=4):
    """Closed-loop throughput of the continuous-batching server (FP16 storage) against plain
    per-call fbp (FP32 and FP16 storage, one synchronised call per job), then an open-loop run
    with Poisson arrivals at 70 % of the measured server capacity for p50 / p99 latency.
    Also reported: the drawn mix against the recipe's expectation, per-class latency, and the
    bound on any batching speed-up over plain calls: the plain calls' wall time over their
    device time (batching can remove launch and synchronisation gaps, not device work)."""
    from curadon_amd.serving import FBPJob, FBPServer
    rng = np.random.default_rng(seed)
    classes = [("fan", n) for n in SERVING_FAN] + [("cone", n) for n in SERVING_CONE]
    prob = [0.7 / len(SERVING_FAN)] * len(SERVING_FAN) + [0.3 / len(SERVING_CONE)] * len(SERVING_CONE)
    draw = rng.choice(len(classes), size=n_jobs, p=prob)
    data = {}
    for ci in sorted(set(draw.tolist())):
        kind, n = classes[ci]
        g32 = serving_geometry(cm, np, kind, n, 32)
        g16 = serving_geometry(cm, np, kind, n, torch.float16)
        if kind == "fan":
            x = torch.tensor(phantom.ellipses_2d(n, seed=ci), dtype=torch.float32, device="cuda")
        else:
            x = torch.tensor(phantom.ellipsoids_3d(n, seed=ci), device="cuda")
        raw = cm.forward(x, g32)
        del x
        data[ci] = (kind, n, g32, g16, raw)
    torch.cuda.synchronize()
    log(f"bench: serving inputs ready ({n_jobs} jobs, {len(data)} classes)")
    tags = {ci: f"{data[ci][0]}{data[ci][1]}" for ci in data}

    def jobs(arrivals=None):
        return [FBPJob(geom=data[c][3], projections=data[c][4], tag=tags[c],
                       arrival=0.0 if arrivals is None else float(arrivals[i]))
                for i, c in enumerate(draw.tolist())]

    # warm-up: FFT plans and kernels of every class, both paths
    for c in data:
        cm.fbp(data[c][4], data[c][2])
        cm.fbp(data[c][4], data[c][3])
    FBPServer().run([FBPJob(geom=data[c][3], projections=data[c][4]) for c in data])
    torch.cuda.synchronize()

    # plain per-call cost per class: device time (events on the current stream) and wall time
    per_class = {}
    stream = torch.cuda.current_stream()
    for c in data:
        dev_ms, wall_ms = [], []
        for _ in range(3):
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            t0 = time.perf_counter()
            e0.record(stream)
            cm.fbp(data[c][4], data[c][3])
            e1.record(stream)
            torch.cuda.synchronize()
            wall_ms.append((time.perf_counter() - t0) * 1e3)
            dev_ms.append(e0.elapsed_time(e1))
        per_class[tags[c]] = {"plain_f16_device_ms": float(np.median(dev_ms)),
                              "plain_f16_wall_ms": float(np.median(wall_ms))}

    def plain(which):
        t0 = time.perf_counter()
        for c in draw.tolist():
            cm.fbp(data[c][4], data[c][which])
            torch.cuda.synchronize()
        return time.perf_counter() - t0

    t_plain32 = plain(2)
    t_plain16 = plain(3)
    closed = jobs()
    srv = FBPServer(hbm_budget=96 << 30)
    st_closed = srv.run(closed)
    log(f"bench: serving closed loop {st_closed['makespan_s']:.3f} s in {srv.rounds} launches; "
        f"plain FP32 {t_plain32:.3f} s, plain FP16 {t_plain16:.3f} s")
    lam = 0.7 * st_closed["jobs_per_s"]
    arrivals = np.cumsum(rng.exponential(1.0 / lam, size=n_jobs))
    arrivals -= arrivals[0]
    opened = jobs(arrivals)
    srv2 = FBPServer(hbm_budget=96 << 30)
    st_open = srv2.run(opened)
    lat = {}
    for j in opened:
        lat.setdefault(j.tag, []).append(j.latency * 1e3)
    # storage error of each class: served FP16 volume against the plain FP32 operator chain
    # (the FP64 pipeline comparison of every class is a GPU test, test_gpu_serving.py)
    errors = {}
    for j, c in zip(closed, draw.tolist()):
        if j.tag in errors:
            continue
        ref = cm.fbp(data[c][4], data[c][2]).double()
        errors[j.tag] = float(torch.linalg.vector_norm(j.volume.double() - ref) /
                              torch.linalg.vector_norm(ref))
    log(f"bench: serving open loop p50 {st_open['latency_ms_p50']:.2f} ms, "
        f"p99 {st_open['latency_ms_p99']:.2f} ms at {lam:.1f} jobs/s")
    counts = {tags[c]: int((draw == c).sum()) for c in data}
    expected = {f"{k}{n}": n_jobs * p for (k, n), p in zip(classes, prob)}
    dev_sum = sum(per_class[tags[c]]["plain_f16_device_ms"] for c in draw.tolist()) * 1e-3
    wall_sum = sum(per_class[tags[c]]["plain_f16_wall_ms"] for c in draw.tolist()) * 1e-3
    for t, v in per_class.items():
        v["jobs_drawn"] = counts.get(t, 0)
        v["jobs_expected"] = expected.get(t, 0.0)
        if t in lat:
            v["open_loop_p50_ms"] = float(np.percentile(lat[t], 50))
            v["open_loop_p99_ms"] = float(np.percentile(lat[t], 99))
            v["open_loop_max_ms"] = float(np.max(lat[t]))
        v["rel_l2_vs_plain_f32"] = errors.get(t)
    return {
        "workload": "FBP/FDK job mix: 70% fan (n in 128..2048, D = ceil(sqrt2 n)), 30% cone "
                    "(n in 128..384, panel 1.5n); A = 2n over 2pi, DSO 2n, DSD 3n; FP32 raw "
                    "projections resident in HBM; cosine weight + ramp filter + weighted "
                    "backprojection",
        "jobs": n_jobs,
        "storage": "f16 (filtered projections and volume), FP32 filtering and accumulation",
        "closed_loop": {
            "serving_jobs_per_s": st_closed["jobs_per_s"],
            "plain_f32_jobs_per_s": n_jobs / t_plain32,
            "plain_f16_jobs_per_s": n_jobs / t_plain16,
            "speedup_vs_plain_f32": t_plain32 / st_closed["makespan_s"],
            "speedup_vs_plain_f16": t_plain16 / st_closed["makespan_s"],
            "speedup_bound_vs_plain_f16": wall_sum / dev_sum,
            "plain_f16_device_s": dev_sum,
            "launches": srv.rounds,
        },
        "open_loop": dict(st_open, arrival_rate_per_s=lam, launches=srv2.rounds,
                          max_active=srv2.max_active_seen),
        "per_class": per_class,
    }


def kernels_only(torch, cm, phantom, np, which, reps):
    """The operators of one workload, `reps` times each after one untimed call, for a profiler
    (rocprofv3 counters and traces attribute each kernel by name)."""
    if which == "fan2d":
        geom, x, sino, vol = build_case(torch, cm, phantom, np)
        ops = [("forward", lambda: cm.forward(x, sino, geom)),
               ("backprojection", lambda: cm.backward(sino, vol, geom)),
               ("adjoint", lambda: cm.adjoint(sino, vol, geom)),
               ("backprojection transpose", lambda: cm.backward_transpose(x, geom, sino))]
    else:
        geom = cm.ConeGeometry(DSD=768.0, DSO=512.0,
                               angles=np.linspace(0, 2 * np.pi, 360, endpoint=False),
                               vol_shape=(256, 256, 256), det_shape=(384, 384))
        x = torch.tensor(phantom.ellipsoids_3d(256, seed=0), device="cuda:0")
        sino = torch.empty(360, 384, 384, device="cuda:0")
        vol = torch.zeros_like(x)
        ops = [("forward", lambda: cm.forward(x, sino, geom)),
               ("backprojection", lambda: cm.backward(sino, vol, geom)),
               ("adjoint", lambda: cm.adjoint(sino, vol, geom)),
               ("backprojection transpose", lambda: cm.backward_transpose(x, geom, sino))]
    geom.plan
    torch.cuda.synchronize()
    for name, fn in ops:
        for _ in range(reps + 1):
            fn()
        torch.cuda.synchronize()
        log(f"bench: {which} {name} x {reps + 1} done")


def copy_calibration(torch, mib, reps):
    """Plain device-to-device copies of `mib` MiB, for scaling HBM byte counters."""
    n = mib * (1 << 20) // 4
    a = torch.ones(n, device="cuda:0")
    b = torch.empty_like(a)
    for _ in range(reps):
        b.copy_(a)
    torch.cuda.synchronize()
    log(f"bench: {reps} copies of {mib} MiB done")


def cpu_baseline():
    """The oracle's FP32 port of the reference arithmetic on this host's cores
    (oracle/cpu_baseline.py): 64^2 with 3 warm-ups + 50 runs of the forward and of the
    backprojection, 1024^2 with 1 warm-up + 5 full runs, a stated sample of the 256^3 cone beam,
    and the 64^2 protocol on one thread."""
    from oracle import cpu_baseline as cb
    r = cb.run(full_cone=False)
    f = r["fan1024"]
    return {
        "value": f["pairs_per_s"], "unit": "fwd+bwd pairs/s", "cores": r["threads"],
        "kind": "port",
        "sample": "FP32 port of the reference arithmetic (oracle/cpu_baseline.py) on the 1024^2 "
                  "workload: 1 warm-up + 5 full runs of the forward and of the backprojection "
                  "on all host threads",
        "cpu_model": r["cpu_model"],
        "fan1024": f, "fan64": r["fan64"], "cone256": r["cone256"],
        "single_thread": r["single_thread"],
        "full_runs": full_runs(),
    }


def full_runs():
    """The committed full-operator runs of the same port (`oracle/cpu_baseline.py --full-cone`,
    1 + 5 full 256^3 cone-beam runs and one full 1024^2 run on one thread; too long for every
    bench run), or None."""
    try:
        with open(os.path.join(ROOT, "profiles", "cpu_baseline_full.json")) as fh:
            d = json.load(fh)
    except (OSError, ValueError):
        return None
    return {"source": "profiles/cpu_baseline_full.json (committed measurement on the GPU host)",
            "threads": d.get("threads"), "cone256": d.get("cone256"),
            "single_thread_fan1024": d.get("single_thread", {}).get("fan1024")}


def roofline(kernel, op_ms, tap_bytes):
    """Rooflines of `kernel` (fwd2d or bp2d) at the live mean op time `op_ms`.

    Per-launch counts come from the committed rocprofv3 counter summary of this workload
    (profiles/counters_fan2d_1024.json): VALU wave instructions (SQ_INSTS_VALU, 64 lanes each),
    LDS bytes (SQ_INSTS_LDS_{LOAD,STORE,ATOMIC}_BANDWIDTH, 64-byte units) and HBM bytes
    (FETCH_SIZE and WRITE_SIZE scaled by a plain copy). Each is divided by the op time measured
    here with HIP events on the launch stream (the op includes its staging kernel, about 10 us)
    and set against the MI355X peak; `bound` names the roof closest to binding. The tap count
    (every interpolation tap priced as an HBM read) is kept as an equivalent figure only."""
    t = op_ms * 1e-3
    tap = {"GB_per_s": tap_bytes / t / 1e9, "bytes_per_launch": tap_bytes,
           "note": "every interpolation tap counted as a 4-byte HBM read; LDS, L1 and L2 serve "
                   "almost all of them, so this is an equivalent rate, not a bound"}
    try:
        with open(COUNTERS_FILE) as f:
            k = json.load(f)["kernels"][kernel]
    except (OSError, KeyError, ValueError):
        return {"bound": None, "kernel": kernel, "achieved": None, "peak": None, "unit": None,
                "frac": None, "traffic": None, "equivalent": tap,
                "note": "no committed counter summary for this kernel"}
    d = k["derived"]
    roofs = {
        "valu": {"achieved": d["valu_wave_instr_per_launch"] * 64 / t / 1e12,
                 "peak": VALU_PEAK_LANES / 1e12, "unit": "T FP32 lane-instr/s"},
        "lds": {"achieved": d["lds_bytes_per_launch"] / t / 1e9, "peak": LDS_PEAK_BPS / 1e9,
                "unit": "GB/s"},
        "hbm": {"achieved": d["hbm_bytes_per_launch"] / t / 1e9, "peak": HBM_PEAK_BPS / 1e9,
                "unit": "GB/s"},
    }
    for r in roofs.values():
        r["frac"] = r["achieved"] / r["peak"]
    bound = max(roofs, key=lambda n: roofs[n]["frac"])
    b = roofs[bound]
    return {"bound": bound, "kernel": k["kernel"], "achieved": b["achieved"], "peak": b["peak"],
            "unit": b["unit"], "frac": b["frac"], "traffic": d["hbm_bytes_per_launch"],
            "roofs": roofs, "lds_bank_conflict_frac": d["lds_bank_conflict_frac"],
            "l2_hit": d["l2_hit"], "equivalent": tap,
            "source": "per-launch counts: profiles/counters_fan2d_1024.json; time: HIP events "
                      "around the op on its launch stream, mean over the timed steps"}


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-extra", action="store_true",
                    help="skip the secondary workloads (256^3 cone beam, batched 512^2)")
    ap.add_argument("--serving-jobs", type=int, default=400,
                    help="jobs in the FBP serving workload (0 skips it)")
    ap.add_argument("--serving-only", action="store_true",
                    help="run only the FBP serving workload and print its JSON")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run plain device copies of MIB MiB (HBM counter calibration)")
    ap.add_argument("--kernels-only", choices=("fan2d", "cone3d"), default=None,
                    help="only run the forward, backprojection and adjoint of the 1024^2 fan "
                         "beam or the 256^3 cone beam, --steps times each (for profilers)")
    args = ap.parse_args()
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print("bench.py measures one GPU only: the projectors run on a single MI355X here and "
              "no multi-GPU figure is produced (run with --gpus 1).", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    import curadon_amd as cm
    from curadon_amd import phantom

    if args.copy_calibration:
        copy_calibration(torch, args.copy_calibration, max(1, args.steps))
        return
    if args.kernels_only:
        kernels_only(torch, cm, phantom, np, args.kernels_only, max(1, args.steps))
        return
    if args.serving_only:
        print(json.dumps(serving_case(torch, cm, phantom, np, max(1, args.serving_jobs))),
              flush=True)
        return

    log("bench: building plan and inputs")
    geom, img, sino, vol = build_case(torch, cm, phantom, np)
    info = geom.plan.info()
    fwd_bytes, bwd_bytes = tap_bytes(info, WORKLOAD)
    stream = torch.cuda.current_stream()

    for _ in range(args.warmup):
        cm.forward(img, sino, geom)
        cm.backward(sino, vol, geom)
    torch.cuda.synchronize()
    log(f"bench: warm-up done ({args.warmup} steps), timing {args.steps} steps")

    ev = [[torch.cuda.Event(enable_timing=True) for _ in range(3)] for _ in range(args.steps)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(args.steps):
        ev[k][0].record(stream)
        cm.forward(img, sino, geom)
        ev[k][1].record(stream)
        cm.backward(sino, vol, geom)
        ev[k][2].record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    fwd_ms = np.array([e[0].elapsed_time(e[1]) for e in ev])
    bwd_ms = np.array([e[1].elapsed_time(e[2]) for e in ev])
    step_ms = fwd_ms + bwd_ms
    log(f"bench: timed region {wall:.3f} s; fwd {fwd_ms.mean():.3f} ms, bwd {bwd_ms.mean():.3f} ms")

    fwd_gbs = fwd_bytes / (fwd_ms.mean() * 1e-3) / 1e9
    bwd_gbs = bwd_bytes / (bwd_ms.mean() * 1e-3) / 1e9
    dominant = "fwd2d" if fwd_ms.mean() >= bwd_ms.mean() else "bp2d"
    roof = (roofline("fwd2d", float(fwd_ms.mean()), fwd_bytes) if dominant == "fwd2d" else
            roofline("bp2d", float(bwd_ms.mean()), bwd_bytes))

    # matched adjoint (autograd's backward of the forward), timed separately
    adj_vol = torch.zeros_like(vol)
    cm.adjoint(sino, adj_vol, geom)
    torch.cuda.synchronize()
    ea, eb = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    n_adj = max(1, min(5, args.steps))
    ea.record(stream)
    for _ in range(n_adj):
        cm.adjoint(sino, adj_vol, geom)
    eb.record(stream)
    torch.cuda.synchronize()
    adj_ms = ea.elapsed_time(eb) / n_adj
    log(f"bench: adjoint {adj_ms:.3f} ms")
    # transpose of the backprojection (TorchBackprojection's gradient), timed separately
    bt_sino = torch.empty_like(sino)
    cm.backward_transpose(vol, geom, bt_sino)
    torch.cuda.synchronize()
    ea.record(stream)
    for _ in range(n_adj):
        cm.backward_transpose(vol, geom, bt_sino)
    eb.record(stream)
    torch.cuda.synchronize()
    bt_ms = ea.elapsed_time(eb) / n_adj
    log(f"bench: backprojection transpose {bt_ms:.3f} ms")

    extra = None
    if not args.no_extra:
        log("bench: secondary workloads")
        extra = {"cone3d": cone_case(torch, cm, phantom, np, stream, 3),
                 "batch2d": batch_case(torch, cm, phantom, np, stream, 3)}
        log(f"bench: cone3d {extra['cone3d']['fwd3d']['ms']:.2f} / {extra['cone3d']['bp3d']['ms']:.2f} ms")

    cold = None
    if not args.no_extra:
        log("bench: cold calls")
        cold = cold_calls(torch, cm, phantom, np)

    serving = None
    if args.serving_jobs > 0:
        log("bench: FBP serving workload")
        serving = serving_case(torch, cm, phantom, np, args.serving_jobs)

    cpu = None
    if not args.no_cpu_baseline:
        log("bench: CPU baseline (oracle FP32 port)")
        cpu = cpu_baseline()

    props = torch.cuda.get_device_properties(0)
    pair_ms = wall / args.steps * 1e3
    out = {
        "metric": "fwd+bwd projections/s (2D fan, FP32)",
        "value": args.steps / wall,
        "unit": "fwd+bwd pairs/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": pair_ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded random-ellipse phantom)",
        "config": {
            "workload": "fan2d 1024x1024 f32, 1449 cells, 720 angles over 2pi, DSO 2048, "
                        "DSD 3072; step = forward + voxel-driven backprojection",
            "batch": 1,
        },
        "ops": {
            "fwd2d": {"ms_mean": float(fwd_ms.mean()), "ms_p50": float(np.percentile(fwd_ms, 50)),
                      "ms_p99": float(np.percentile(fwd_ms, 99)), "op_per_s": 1e3 / fwd_ms.mean(),
                      "tap_equiv_GB_per_s": fwd_gbs, "tap_bytes": fwd_bytes},
            "bp2d": {"ms_mean": float(bwd_ms.mean()), "ms_p50": float(np.percentile(bwd_ms, 50)),
                     "ms_p99": float(np.percentile(bwd_ms, 99)), "op_per_s": 1e3 / bwd_ms.mean(),
                     "tap_equiv_GB_per_s": bwd_gbs, "tap_bytes": bwd_bytes},
            "adj2d": {"ms_mean": adj_ms, "op_per_s": 1e3 / adj_ms},
            "bp2d_transpose": {"ms_mean": bt_ms, "op_per_s": 1e3 / bt_ms},
            "pair": {"ms_p50": float(np.percentile(step_ms, 50)),
                     "ms_p99": float(np.percentile(step_ms, 99))},
            "total_samples": info["total_samples"],
            # plan on a fresh geometry plus its first forward and backprojection
            "cold_ms": (cold or {}).get("fan2d_1024", {}).get("plan_and_first_pair_ms"),
        },
        "roofline": roof,
        "cpu_baseline": cpu,
        "cold": cold,
        "extra": extra,
        "serving": serving,
        "device": {"name": props.name, "cus": props.multi_processor_count,
                   "gcn_arch": getattr(props, "gcnArchName", "")},
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
