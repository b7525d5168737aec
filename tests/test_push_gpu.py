"""Parity of the HIP batch push-cutter with the FP64 oracle (GPU).

By default the device pushes exactly the triangles the reference's KD search
returns for each fiber (a superset of the box overlaps; the reference pushes
all of them, batchpushcutter.cpp:170-181) and merges the pair intervals like
Fiber::addInterval (fiber.cpp:58-93); it is compared with the oracle's KD
search (search_mode 0).  With the exact box filter option it is compared
with the oracle's exact filter (search_mode 2).  Bar: per fiber the same
number of intervals, the same endpoints (bit for bit; BullCutter 1e-9 for the
device atan/sin inside the offset-ellipse solve), the same CC types and CC
points.  Interval order within a fiber is compared after sorting by the lower
end (the device returns them sorted; the reference's order is its merge
history).
"""
import math

import numpy as np
import pytest

from cutters import ALL, BULL10, CONE10
from opencamlib_amd import capi, workloads as W
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def fibers_for(tris, R, zs, n, direction):
    out = []
    for z in zs:
        xf, yf = W.waterline_fibers(tris, R, z, n)
        out.append(xf if direction == "x" else yf)
    return np.concatenate(out)


def split(res, nf):
    offs = res["offsets"]
    return [(res["lohi"][offs[i]:offs[i + 1]], res["cc"][offs[i]:offs[i + 1]]) for i in range(nf)]


def compare_push(tris, spec, fib, direction, label, exact_filter=False, dev=None, precision="exact64"):
    """precision "exact64" pushes every candidate pair; the other modes leave
    out the pairs that cannot change their fiber (same results)."""
    search_mode = 2 if exact_filter else 0
    own = dev is None
    if own:
        dev = capi.DeviceMesh(tris)
    got = dev.push(spec, fib, direction, precision=precision, exact_filter=exact_filter)
    if own:
        dev.close()
    ref = O.Mesh(tris).batch_push(O.cutter_from_spec(spec), fib, 0 if direction == "x" else 1,
                                  search_mode=search_mode)
    nf = len(fib)
    g, r = split(got, nf), split(ref, nf)
    tol = 1e-9 if spec["kind"] in ("bull", "composite") else 0.0
    bad_count = bad_end = bad_cc = 0
    worst = 0.0
    for i in range(nf):
        (gl, gc), (rl, rc) = g[i], r[i]
        if len(gl) != len(rl):
            bad_count += 1
            continue
        o = np.argsort(rl[:, 0], kind="stable")
        rl, rc = rl[o], rc[o]
        if len(gl):
            d = np.max(np.abs(gl - rl))
            worst = max(worst, d)
            if d > tol:
                bad_end += 1
            if not np.array_equal(gc[:, [3, 7]], rc[:, [3, 7]]) or np.max(np.abs(gc - rc)) > max(tol, 1e-12) * 1e3:
                bad_cc += 1
    print(f"[push] {label} {spec['kind']} {precision} dir={direction}: fibers={nf} intervals={len(got['lohi'])} "
          f"count_mismatch={bad_count} endpoint_mismatch={bad_end} cc_mismatch={bad_cc} max|dt|={worst:.3e} "
          f"stats={got['stats']}", flush=True)
    if precision == "exact64":
        assert got["stats"]["skipped"] == 0
    compare_push.last_stats = got["stats"]
    return bad_count, bad_end, bad_cc


PRECISIONS = ["exact64", "mixed"]


@pytest.fixture(scope="module")
def small_scene():
    return W.spheres_scene(4, 2, 60.0, seed=5)


@pytest.mark.parametrize("precision", PRECISIONS)
@pytest.mark.parametrize("direction", ["x", "y"])
@pytest.mark.parametrize("name", ["cyl", "ball", "bull", "cone", "cylball"])
def test_push_small_scene(gpu, small_scene, name, direction, precision):
    spec = ALL[name]
    R = max(spec.get("d", 10.0), 10.0) / 2
    fib = fibers_for(small_scene, R, [1.0, 4.5, 9.0, 17.0], 60, direction)
    c, e, k = compare_push(small_scene, spec, fib, direction, "small", precision=precision)
    assert c == 0 and e == 0 and k == 0


@pytest.mark.parametrize("precision", PRECISIONS)
@pytest.mark.parametrize("spec", [BULL10, CONE10], ids=["bull", "cone"])
@pytest.mark.parametrize("direction", ["x", "y"])
def test_push_config2_subset(gpu, spec, direction, precision):
    """The 204,802-triangle sphere scene; 300 fibers per direction at three
    of the twenty waterline levels.  In mixed precision most candidate
    pairs cannot change their fiber and are left out."""
    tris = W.config2_mesh()
    levels = W.config2_levels()
    fib = fibers_for(tris, spec["d"] / 2, [levels[0], levels[7], levels[15]], 300, direction)
    fib = fib[::3]
    c, e, k = compare_push(tris, spec, fib, direction, "C2", precision=precision)
    assert c == 0 and e == 0 and k == 0
    st = compare_push.last_stats
    if precision != "exact64":
        dev = capi.DeviceMesh(tris)
        every = dev.push(spec, fib, direction, precision="exact64")["stats"]
        dev.close()
        assert st["pairs"] < 0.5 * every["pairs"], (st, every)


@pytest.mark.parametrize("precision", PRECISIONS)
@pytest.mark.parametrize("name", ["bull", "cone"])
def test_push_exact_filter_option(gpu, small_scene, name, precision):
    """Only pairs whose search boxes overlap: same as the oracle's exact
    filter.  On this scene the KD superset adds intervals (shaft contacts of
    triangles above f.z + L), so the two options differ."""
    spec = ALL[name]
    fib = fibers_for(small_scene, 5.0, [2.0, 6.0], 60, "x")
    c, e, k = compare_push(small_scene, spec, fib, "x", "exact-filter", exact_filter=True, precision=precision)
    assert c == 0 and e == 0 and k == 0


def test_push_errors_and_empty(gpu):
    tris = W.heightfield(4, 10.0, bumps=1, seed=1)
    dev = capi.DeviceMesh(tris)
    bad = np.array([[0, 1, 0, 10, 2, 0]], dtype=float)  # not along X
    with pytest.raises(capi.OclError, match="axis"):
        dev.push(BULL10, bad, "x")
    tilted = np.array([[0, 1, 0, 10, 1, 1]], dtype=float)
    with pytest.raises(capi.OclError, match="horizontal"):
        dev.push(BULL10, tilted, "x")
    res = dev.push(BULL10, np.zeros((0, 6)), "x")
    assert len(res["lohi"]) == 0
    far = np.array([[0, 500, 0, 10, 500, 0]], dtype=float)
    res = dev.push(BULL10, far, "x")
    assert list(res["offsets"]) == [0, 0]
    dev.close()
    empty = capi.DeviceMesh(np.zeros((0, 9)))
    res = empty.push(BULL10, np.array([[0, 1, 0, 10, 1, 0]], dtype=float), "x")
    assert list(res["offsets"]) == [0, 0]
    empty.close()


def comb(n, pitch):
    """n thin tent ridges (base 0.2 mm, height 5 mm) along y at x = pitch k."""
    tris = []
    for k in range(n):
        x = pitch * k
        a, b, c, d = (x - 0.1, -10, 0), (x + 0.1, -10, 0), (x + 0.1, 10, 0), (x - 0.1, 10, 0)
        t1, t2 = (x, -10, 5), (x, 10, 5)
        tris += [a + d + t2, a + t2 + t1, b + t1 + t2, b + t2 + c]
    return np.array(tris, dtype=float)


@pytest.mark.parametrize("precision", PRECISIONS)
@pytest.mark.parametrize("direction", ["x", "y"])
@pytest.mark.parametrize("kind", ["bull", "cone"])
def test_push_comb_overflow_and_capacity(gpu, kind, direction, precision):
    """Fibers across 40 ridges hold 40 disjoint intervals (the sparse ones
    2): lists that fit the LDS list (64 entries for the BullCutter, 128 for
    the others) and go to the compact pool.  A comb of 150 ridges gives
    150-interval lists that outgrow LDS and run again in the retry pass,
    which merges in global memory.  Pushes in a row
    on one mesh (dense, sparse, dense again, longer) all match the oracle."""
    spec = {"bull": {"kind": "bull", "d": 2.0, "r": 0.5, "L": 10.0},
            "cone": {"kind": "cone", "d": 2.0, "angle": math.pi / 4, "L": 10.0}}[kind]
    def mesh(n):
        t = comb(n, 3.0)
        if direction == "y":
            t = t.reshape(-1, 3, 3)[:, :, [1, 0, 2]].reshape(-1, 9)
        return t
    tris = mesh(40)
    def fibers(x1):
        rows = []
        for z in (1.0, 3.0):
            for a in (-3.0, 0.0, 4.0):
                rows.append([-5, a, z, x1, a, z] if direction == "x" else [a, -5, z, a, x1, z])
        return np.array(rows, dtype=float)
    dense, sparse = fibers(125.0), fibers(8.0)
    dev = capi.DeviceMesh(tris)
    for label, fib, nmin in (("comb-dense", dense, 40), ("comb-sparse", sparse, 2), ("comb-dense-again", dense, 40)):
        got = dev.push(spec, fib, direction, precision=precision)
        longest = int(np.max(np.diff(got["offsets"])))
        assert longest >= nmin, (label, longest)
        c, e, k = compare_push(tris, spec, fib, direction, label, dev=dev, precision=precision)
        assert c == 0 and e == 0 and k == 0, label
    dev.close()
    tris = mesh(150)
    dev = capi.DeviceMesh(tris)
    fib = fibers(455.0)
    got = dev.push(spec, fib, direction, precision=precision)
    assert int(np.max(np.diff(got["offsets"]))) >= 150
    c, e, k = compare_push(tris, spec, fib, direction, "comb-150", dev=dev, precision=precision)
    assert c == 0 and e == 0 and k == 0
    dev.close()
    # 2,000 fibers of 40 intervals: 80,000 entries, more than the first
    # pool (4 per fiber), so most lists go to the retry pass; the next push
    # sizes its pool from this one's total
    tris = mesh(40)
    dev = capi.DeviceMesh(tris)
    a = np.linspace(-9.0, 9.0, 2000)
    z = np.full_like(a, 1.0)
    if direction == "x":
        fib = np.stack([np.full_like(a, -5.0), a, z, np.full_like(a, 125.0), a, z], axis=1)
    else:
        fib = np.stack([a, np.full_like(a, -5.0), z, a, np.full_like(a, 125.0), z], axis=1)
    for label in ("pool-full", "pool-sized"):
        c, e, k = compare_push(tris, spec, fib, direction, label, dev=dev, precision=precision)
        assert c == 0 and e == 0 and k == 0, label
    dev.close()
