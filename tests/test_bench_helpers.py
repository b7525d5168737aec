"""CPU checks of bench.py's figures that do not need a GPU."""
import importlib.util
import os

import numpy as np

from opencamlib_amd import workloads as W

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _bench():
    spec = importlib.util.spec_from_file_location("bench", os.path.join(ROOT, "bench.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def test_overlap_pairs_matches_brute_force():
    # the reference's overlaps() (millingcutter.cpp:422-433) evaluated pair by pair
    b = _bench()
    tris = W.heightfield(16, 20.0, bumps=2, steps=1, seed=11)
    cl = W.cl_grid(40, 40, 0.5, origin=(-1.0, 0.5))
    t = tris.reshape(-1, 3, 3)
    for R in (0.5, 2.0, 3.0):
        want = 0
        for x, y, _ in cl:
            off = ((t[:, :, 0].max(1) < x - R) | (t[:, :, 0].min(1) > x + R) |
                   (t[:, :, 1].max(1) < y - R) | (t[:, :, 1].min(1) > y + R))
            want += int(np.sum(~off))
        assert b.overlap_pairs(tris, cl, R) == want


def test_overlap_pairs_needs_a_product_grid():
    b = _bench()
    tris = W.heightfield(4, 10.0, bumps=1, seed=1)
    cl = W.cl_random(100, 10.0, seed=3)
    assert b.overlap_pairs(tris, cl, 1.0) is None
