"""The roofline that bench.py reports, checked on the CPU against the committed counter summary
(profiles/counters_fan2d_1024.json): every roof's fraction is the per-launch count over the op
time against the MI355X peak, no roof is passed at the kernel's own traced duration, and
`bound` names the roof closest to binding."""
import json
import os

import pytest

import bench

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def counters():
    with open(os.path.join(ROOT, "profiles", "counters_fan2d_1024.json")) as f:
        return json.load(f)


@pytest.mark.parametrize("kernel", ["fwd2d", "bp2d"])
def test_roofs_at_the_traced_duration(kernel):
    k = counters()["kernels"][kernel]
    d = k["derived"]
    t_ms = k["avg_duration_us"] / 1e3
    r = bench.roofline(kernel, t_ms, 1)
    assert r["kernel"] == k["kernel"]
    roofs = r["roofs"]
    # each fraction is the count over the time against the peak
    t = t_ms * 1e-3
    assert roofs["valu"]["frac"] == pytest.approx(
        d["valu_wave_instr_per_launch"] * 64 / t / (128 * 256 * 2.4e9), rel=1e-12)
    assert roofs["lds"]["frac"] == pytest.approx(
        d["lds_bytes_per_launch"] / t / (128 * 256 * 2.4e9), rel=1e-12)
    assert roofs["hbm"]["frac"] == pytest.approx(d["hbm_bytes_per_launch"] / t / 8e12, rel=1e-12)
    # at the duration the counters were taken with, the summary's own fractions come back
    assert roofs["valu"]["frac"] == pytest.approx(d["valu_frac_of_peak"], rel=1e-6)
    assert roofs["lds"]["frac"] == pytest.approx(d["lds_frac_of_peak"], rel=1e-6)
    assert roofs["hbm"]["frac"] == pytest.approx(d["hbm_frac_of_peak"], rel=1e-6)
    # no roof is exceeded, and the bound is the largest fraction
    assert all(0.0 < v["frac"] <= 1.0 for v in roofs.values())
    assert r["bound"] == max(roofs, key=lambda n: roofs[n]["frac"])
    assert r["frac"] == roofs[r["bound"]]["frac"]
    assert r["traffic"] == d["hbm_bytes_per_launch"]


def test_headline_kernel_is_not_hbm_bound():
    """The forward's traffic beyond L2 (FETCH_SIZE and WRITE_SIZE, scaled by a plain copy) stays
    well below the HBM roof at its traced duration, so the roof that binds is LDS or VALU issue.
    Those bytes are mostly the staged bands' re-reads of the two image copies (8.4 MB together,
    more than one XCD's 4 MB L2) and the 33 MB per-ray table."""
    k = counters()["kernels"]["fwd2d"]
    r = bench.roofline("fwd2d", k["avg_duration_us"] / 1e3, 1)
    assert r["roofs"]["hbm"]["frac"] < 0.5
    assert r["bound"] in ("lds", "valu")


def test_missing_kernel_reports_no_roofline():
    r = bench.roofline("no_such_kernel", 1.0, 4)
    assert r["bound"] is None and r["frac"] is None
    assert r["equivalent"]["GB_per_s"] == pytest.approx(4 / 1e-3 / 1e9)
