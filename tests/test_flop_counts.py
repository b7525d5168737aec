"""The algorithmic flop counts per evaluated pair that bench.py's roofline
lines use, counted from the kernels' own contact code (CPU).

tests/flops/flop_drop.cpp runs the FP32 selection's contact function on the
host with a counting float type: it has no data-dependent branches, so one
pair gives the exact count (add/sub/mul = 1, fma = 2, sqrt/rsqrt/rcp/div = 1,
compare/select/min/max = 0).  tests/flops/flop_push.cpp compiles the FP64
push contact code with `double` replaced by a counting type and runs it over
the candidate pairs (the reference's KD-search superset, from the oracle) of
40 seeded X fibers of the push workload; the push branches on the geometry,
so its count is that sample's average, Brent iterations of the BullCutter
edge solves included."""
import ctypes as C
import os
import subprocess
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
FLOPS = os.path.join(ROOT, "tests", "flops")
BUILD = os.path.join(FLOPS, "_build")
sys.path.insert(0, ROOT)
import bench  # noqa: E402  (the constants under test)


def build(name):
    os.makedirs(BUILD, exist_ok=True)
    out = os.path.join(BUILD, f"lib{name}.so")
    subprocess.run(["g++", "-O1", "-std=c++17", "-fPIC", "-shared", "-fvisibility=hidden", "-o", out,
                    os.path.join(FLOPS, f"{name}.cpp")], check=True, capture_output=True)
    return C.CDLL(out)


def flops(c):
    add, mul, fma, sq, div = c[:5]
    other = c[5] if len(c) > 6 else 0
    return add + mul + 2 * fma + sq + div + other


def test_drop_flops_per_pair_match_the_bench_table():
    L = build("flop_drop")
    got = {}
    for k, name in enumerate(["cyl", "ball", "bull", "cone"]):
        o = (C.c_ulonglong * 5)()
        assert L.fc_drop_pair(k, o) == 0
        got[name] = flops(list(o))
    print("FP32 selection flops per evaluated pair:", got)
    assert got == bench.F_PAIR


PUSH = {"bull": (2, {"kind": "bull", "d": 10.0, "r": 2.0, "L": 20.0}, [10.0, 2.0, 20.0]),
        "cone": (3, {"kind": "cone", "d": 10.0, "angle": np.pi / 4, "L": 10.0}, [10.0, np.pi / 4, 10.0])}


@pytest.mark.parametrize("name", ["bull", "cone"])
def test_push_flops_per_pair_match_the_bench_table(name):
    from opencamlib_amd import workloads as W
    from oracle import oracle as O
    L = build("flop_push")
    kind, spec, par = PUSH[name]
    tris = np.ascontiguousarray(W.config2_mesh())
    om = O.Mesh(tris)
    cut = O.cutter_from_spec(spec)
    fibs = np.concatenate([W.waterline_fibers(tris, spec["d"] / 2, z, 2000)[0] for z in W.config2_levels()])
    idx = np.random.default_rng(5).choice(len(fibs), 40, replace=False)
    pairs = [(j, int(k)) for j, i in enumerate(idx) for k in om.kd_search_push(cut, fibs[i], 0)]
    fs = np.ascontiguousarray(fibs[idx])
    pr = np.ascontiguousarray(np.array(pairs, dtype=np.uint32))
    sp = np.array(par, dtype=np.float64)
    o = (C.c_ulonglong * 7)()
    assert L.fc_push_pairs(kind, sp.ctypes.data_as(C.c_void_p), tris.ctypes.data_as(C.c_void_p),
                           fs.ctypes.data_as(C.c_void_p), pr.ctypes.data_as(C.c_void_p), C.c_size_t(len(pr)), o) == 0
    per_pair = flops(list(o)[:6] + [0]) / len(pr)
    brent = o[6] / len(pr)
    print(f"FP64 push {name}: {len(pr)} pairs, {per_pair:.1f} flops/pair, {brent:.2f} Brent iterations/pair")
    assert abs(per_pair - bench.F_PUSH[name]) <= 0.03 * bench.F_PUSH[name]
    if name == "bull":
        ob = (C.c_ulonglong * 7)()
        assert L.fc_brent_solve(sp.ctypes.data_as(C.c_void_p), tris.ctypes.data_as(C.c_void_p),
                                fs.ctypes.data_as(C.c_void_p), pr.ctypes.data_as(C.c_void_p), C.c_size_t(len(pr)),
                                ob) == 0
        per_it = flops(list(ob)[:6] + [0]) / ob[6]
        print(f"one Brent solve: {ob[6]} iterations, {per_it:.1f} flops/iteration")
        assert 40 <= per_it <= 200


def kept_pairs(tris, spec, fibs, om, cut):
    """The candidate pairs the mixed-precision push keeps: each fiber's
    KD-search candidates in rank order, less those push_bound leaves out (no
    interval possible, or the bound strictly inside the fiber's list as
    built from the batches of 64 kept pairs before), replayed on the host."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from test_push_bound import selcheck
    from opencamlib_amd import capi
    from oracle import oracle as O
    L = selcheck()
    desc = capi.cutter_desc(spec)
    out = np.zeros(2)
    keep = []
    for j, fib in enumerate(fibs):
        fib = np.ascontiguousarray(fib)
        lst, queue = [], []

        def flush():
            nonlocal lst
            for lo, hi in queue:
                nl = []
                for e in lst:
                    if e[1] < lo or e[0] > hi:
                        nl.append(e)
                    else:
                        lo, hi = min(lo, e[0]), max(hi, e[1])
                nl.append((lo, hi))
                lst = sorted(nl)
            queue.clear()

        for k in om.kd_search_push(cut, fib, 0):
            t = np.ascontiguousarray(tris[k])
            r = L.ocl_selcheck_push_bound(t.ctypes.data, C.addressof(desc), fib.ctypes.data, 0, out.ctypes.data)
            if r == 1 or (r == 2 and any(e[0] < out[0] and out[1] < e[1] for e in lst)):
                continue
            keep.append((j, int(k)))
            ne, iv = O.push_triangle(cut, fib, t)
            if ne:
                queue.append((iv[0], iv[1]))
            if len(keep) % 64 == 0:
                flush()
        flush()
    return keep


@pytest.mark.parametrize("name", ["bull", "cone"])
def test_push_flops_per_kept_pair_match_the_bench_table(name):
    """The same count over the pairs the mixed-precision push evaluates
    (bench.F_PUSH_KEPT, the mixed push line's roofline)."""
    from opencamlib_amd import workloads as W
    from oracle import oracle as O
    L = build("flop_push")
    kind, spec, par = PUSH[name]
    tris = np.ascontiguousarray(W.config2_mesh())
    om = O.Mesh(tris)
    cut = O.cutter_from_spec(spec)
    fibs = np.concatenate([W.waterline_fibers(tris, spec["d"] / 2, z, 2000)[0] for z in W.config2_levels()])
    idx = np.random.default_rng(5).choice(len(fibs), 40, replace=False)
    fs = np.ascontiguousarray(fibs[idx])
    pairs = kept_pairs(tris.reshape(-1, 9), spec, fs, om, cut)
    pr = np.ascontiguousarray(np.array(pairs, dtype=np.uint32))
    sp = np.array(par, dtype=np.float64)
    o = (C.c_ulonglong * 7)()
    assert L.fc_push_pairs(kind, sp.ctypes.data_as(C.c_void_p), tris.ctypes.data_as(C.c_void_p),
                           fs.ctypes.data_as(C.c_void_p), pr.ctypes.data_as(C.c_void_p), C.c_size_t(len(pr)), o) == 0
    per_pair = flops(list(o)[:6] + [0]) / len(pr)
    print(f"FP64 push {name}, kept pairs: {len(pr)}, {per_pair:.1f} flops/pair, {o[6] / len(pr):.2f} Brent "
          f"iterations/pair")
    assert abs(per_pair - bench.F_PUSH_KEPT[name]) <= 0.03 * bench.F_PUSH_KEPT[name]
