"""CPU baseline: the oracle's FP32 port of the reference's per-ray / per-voxel projector arithmetic
(orc_*_f32), timed on the host cores. The reference has no CPU path of its own, and its CUDA
sources cannot be built here, so this port is the CPU counterpart the GPU numbers are set beside.
Test and measurement infrastructure only: the product never calls it.

Protocol (the reference's harness times 50 runs after 3 warm-ups, benchmark/bench.py:51-75, and
reports op/s = 1 / mean, :294-296):
  * 64^2 fan beam (91 cells, 360 angles, DSO 128, DSD 192): 3 warm-ups + 50 runs of the forward
    and of the backprojection, separately; mean, p50, p99 and op/s;
  * 1024^2 fan beam (1449 cells, 720 angles, DSO 2048, DSD 3072): 1 warm-up + 5 full runs;
  * 256^3 cone beam (384^2 panel, 360 angles, DSO 512, DSD 768): 1 warm-up + 5 full runs
    (``full_cone``), or a stated sample of projections / slices;
all on all host threads; then the 64^2 protocol on one thread (and, with ``full_cone``, one
full 1024^2 run on one thread).

Usage: python oracle/cpu_baseline.py [--full-cone] [--json OUT]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

from oracle import oracle  # noqa: E402


def _rec2d(n, D, A, dso, dsd):
    from types import SimpleNamespace
    ang = np.linspace(0, 2 * np.pi, A, endpoint=False)
    return SimpleNamespace(vol_shape=np.array([n, n]), vol_spacing=np.ones(2),
                           vol_offset=np.zeros(2), det_shape=D, det_spacing=1.0, det_offset=0.0,
                           angles=ang, DSO=np.full(A, dso), DSD=np.full(A, dsd), nangles=A)


def _rec3d(n, U, A, dso, dsd):
    from types import SimpleNamespace
    ang = np.linspace(0, 2 * np.pi, A, endpoint=False)
    return SimpleNamespace(vol_shape=np.array([n, n, n]), vol_spacing=np.ones(3),
                           vol_offset=np.zeros(3), det_shape=np.array([U, U]),
                           det_spacing=np.ones(2), det_offset=np.zeros(2),
                           det_rotation=np.zeros(3),
                           angles=np.vstack([ang, np.zeros(A), np.zeros(A)]), DSO=dso, DSD=dsd,
                           COR=0.0, nangles=A)


def _stats(ts):
    ts = np.asarray(ts) * 1e3
    return {"ms_mean": float(ts.mean()), "ms_p50": float(np.percentile(ts, 50)),
            "ms_p99": float(np.percentile(ts, 99)), "op_per_s": float(1e3 / ts.mean()),
            "runs": int(ts.size)}


def _progress(msg):
    print(f"cpu_baseline: {msg}", file=sys.stderr, flush=True)


def _time(fn, warm, runs, what=None):
    for _ in range(warm):
        fn()
    out = []
    for k in range(runs):
        t0 = time.perf_counter()
        fn()
        out.append(time.perf_counter() - t0)
        if what and out[-1] > 5.0:
            _progress(f"{what} run {k + 1}/{runs}: {out[-1]:.1f} s")
    return out


def fan(n, D, A, dso, dsd, warm, runs, seed=0):
    from curadon_amd import phantom
    rec = _rec2d(n, D, A, dso, dsd)
    img = phantom.ellipses_2d(n, seed=seed).astype(np.float32)
    sino = oracle.fwd2d_f32(rec, img)
    f = _stats(_time(lambda: oracle.fwd2d_f32(rec, img), warm, runs, f"fan{n} forward"))
    b = _stats(_time(lambda: oracle.bp2d_f32(rec, sino), warm, runs, f"fan{n} backprojection"))
    _progress(f"fan{n} done ({oracle.num_threads()} threads)")
    return {"fwd": f, "bwd": b, "pairs_per_s": 1e3 / (f["ms_mean"] + b["ms_mean"])}


def cone(warm, runs, fwd_angles=None, bp_slices=None, seed=0):
    """256^3 cone beam; fwd_angles / bp_slices = a sample (count) instead of the full operator,
    scaled to it."""
    from curadon_amd import phantom
    rec = _rec3d(256, 384, 360, 512.0, 768.0)
    vol = phantom.ellipsoids_3d(256, seed=seed).astype(np.float32)
    fa = 360 if fwd_angles is None else int(fwd_angles)
    bz = 256 if bp_slices is None else int(bp_slices)
    sino = oracle.fwd3d_f32(rec, vol, angles=(0, fa))
    f = _time(lambda: oracle.fwd3d_f32(rec, vol, angles=(0, fa)), warm, runs, "cone256 forward")
    b = _time(lambda: oracle.bp3d_f32(rec, sino, slabs=(0, bz)), warm, runs,
              "cone256 backprojection")
    _progress("cone256 done")
    f = _stats(np.asarray(f) * 360 / fa)
    b = _stats(np.asarray(b) * 256 / bz)
    return {"fwd": f, "bwd": b, "pairs_per_s": 1e3 / (f["ms_mean"] + b["ms_mean"]),
            "sample": "full operators" if (fa == 360 and bz == 256) else
            f"forward on {fa}/360 projections, backprojection on {bz}/256 slices, scaled"}


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def run(full_cone=False, single_core=True):
    """The protocol above. Returns a dict of the per-workload figures, all threads first."""
    out = {"threads": oracle.num_threads(), "cpu_model": cpu_model()}
    out["fan64"] = fan(64, 91, 360, 128.0, 192.0, 3, 50)
    out["fan1024"] = fan(1024, 1449, 720, 2048.0, 3072.0, 1, 5)
    out["cone256"] = cone(1, 5) if full_cone else cone(1, 2, fwd_angles=8, bp_slices=8)
    if single_core:
        n = oracle.num_threads()
        oracle.set_num_threads(1)
        try:
            out["single_thread"] = {"fan64": fan(64, 91, 360, 128.0, 192.0, 3, 50)}
            if full_cone:
                out["single_thread"]["fan1024"] = fan(1024, 1449, 720, 2048.0, 3072.0, 0, 1)
        finally:
            oracle.set_num_threads(n)
    return out


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--full-cone", action="store_true", help="1 + 5 full 256^3 runs")
    ap.add_argument("--json", default=None, help="write the figures to this file")
    args = ap.parse_args()
    r = run(full_cone=args.full_cone)
    txt = json.dumps(r, indent=1)
    if args.json:
        with open(args.json, "w") as f:
            f.write(txt + "\n")
    print(txt, flush=True)


if __name__ == "__main__":
    main()
