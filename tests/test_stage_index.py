"""Host-side unit test of the staging kernels' flat-index decomposition
(curadon_amd/csrc/stage_index.hpp): the header is compiled with the host C++ compiler, and its
32-bit and 64-bit paths are checked against direct arithmetic, including indices beyond 2^32
that only the 64-bit path handles, and stage_next against decomposing the next index."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "curadon_amd", "csrc")

PROGRAM = r"""
#include "stage_index.hpp"
#include <cstdio>
#include <cstdint>
using namespace cmi;
static uint64_t rng = 0x9e3779b97f4a7c15ull;
static uint64_t next() { rng ^= rng << 13; rng ^= rng >> 7; rng ^= rng << 17; return rng; }
int main() {
    const int shapes[][2] = {{5, 1}, {7, 9}, {581, 580}, {1453, 1}, {65539, 4}, {4, 65539}};
    long bad = 0, checked = 0;
    for (auto &sh : shapes) {
        const int Up = sh[0], Vp = sh[1];
        for (int t = 0; t < 20000; ++t) {
            const bool wide = t & 1;
            const uint64_t i = wide ? next() % (1ull << 40) : next() % (1ull << 32);
            const uint64_t a = i / ((uint64_t)Up * Vp), rem = i % ((uint64_t)Up * Vp);
            const int vp = (int)(rem / Up), up = (int)(rem % Up);
            StageIndex s = stage_index(i, Up, Vp, false);
            bad += s.a != a || s.vp != vp || s.up != up;
            if (!wide) {
                StageIndex n = stage_index(i, Up, Vp, true);
                bad += n.a != a || n.vp != vp || n.up != up;
            }
            stage_next(s, Up, Vp);
            StageIndex q = stage_index(i + 1, Up, Vp, false);
            bad += s.a != q.a || s.vp != q.vp || s.up != q.up;
            ++checked;
        }
    }
    std::printf("checked %ld bad %ld\n", checked, bad);
    return bad != 0;
}
"""


@pytest.mark.skipif(shutil.which("g++") is None, reason="no host C++ compiler")
def test_stage_index_paths(tmp_path):
    src = tmp_path / "t.cpp"
    src.write_text(PROGRAM)
    exe = tmp_path / "t"
    subprocess.run(["g++", "-std=c++17", "-O1", "-I", CSRC, str(src), "-o", str(exe)],
                   check=True)
    r = subprocess.run([str(exe)], capture_output=True, text=True)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "bad 0" in r.stdout
