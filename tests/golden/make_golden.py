"""Writes the golden fixtures under tests/golden/ from the FP64 oracle.

The reference ships no numeric fixtures for its projectors (its tests compare perceptual
hashes against pyelsa / phantominator output, python/tests/test_forward_2d.py:91-138), so these
fixtures are outputs of the oracle, kept to catch regressions of the oracle itself; the
closed-form and library checks in tests/test_oracle_*.py are what tie the oracle to the
reference's arithmetic.

Run from the repository root:  python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from geoms import cone, tiny  # noqa: E402
from oracle import oracle  # noqa: E402


def main():
    oracle.build()
    g = tiny()
    r = np.random.default_rng(20240801)
    image = r.random((8, 8))
    sino_in = r.random((g.nangles, g.det_shape))
    np.savez(os.path.join(HERE, "fan2d_tiny.npz"), angles=g.angles, image=image,
             sino=oracle.fwd2d(g, image), sino_in=sino_in, bp=oracle.bp2d(g, sino_in),
             adj=oracle.adj2d(g, sino_in))
    print("wrote fan2d_tiny.npz")
    c = cone(8, 6, 5, 8, 16.0, 28.0)
    vol = r.random((8, 8, 8))
    proj_in = r.random((8, 5, 6))
    np.savez(os.path.join(HERE, "cone3d_tiny.npz"), angles=c.angles, volume=vol,
             sino=oracle.fwd3d(c, vol), sino_in=proj_in, bp=oracle.bp3d(c, proj_in),
             bp_matched=oracle.bp3d(c, proj_in, matched=True), adj=oracle.adj3d(c, proj_in))
    print("wrote cone3d_tiny.npz")


if __name__ == "__main__":
    main()
