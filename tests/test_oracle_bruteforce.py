"""Pins of the FP64 oracle by brute force over the contact geometry.

Single-triangle drops of every cutter are checked against the supremum of
p.z - h(q) over the triangle (tests/bruteforce.py).  This pins the Bull
offset-ellipse edge solve (bullcutter.cpp:82-105, ellipse.cpp:185-208) and
the Cone edge branches (conecutter.cpp:109-135) independently of their
formulas.  Cone cutters are built with l=0: the reference's rim facet and
hyperbola edge formulas use `length` where the geometry has `center_height`
(conecutter.cpp:92, 122); with l=0 the two coincide, so the brute force pins
the code paths exactly.  The l>0 variant is pinned against these l=0 results
by tests/test_oracle_composite.py, which also pins composite cutters.

Batch drops are checked across the three candidate searches (reference KD
order, exhaustive overlap scan in KD rank order, all triangles), and push
intervals against drop heights (a fiber point collides iff the drop height
there exceeds the fiber z).
"""
import math

import numpy as np
import pytest

from bruteforce import drop_sup, union_sorted
from opencamlib_amd import workloads as W
from oracle import oracle as O

SPECS = {
    "cyl": {"kind": "cyl", "d": 6.0, "L": 20.0},
    "ball": {"kind": "ball", "d": 6.0, "L": 20.0},
    "bull": {"kind": "bull", "d": 6.0, "r": 1.2, "L": 20.0},
    "cone": {"kind": "cone", "d": 6.0, "angle": math.radians(50), "L": 0.0},
    "cone_steep": {"kind": "cone", "d": 6.0, "angle": math.radians(20), "L": 0.0},
}
TYPE_KIND = {1: "vertex", 3: "edge", 14: "facet", 15: "facet", 16: "facet"}


def random_tri(rng):
    c = rng.uniform(-2, 2, 2)
    pts = []
    for _ in range(3):
        xy = c + rng.uniform(-4, 4, 2)
        pts.append([xy[0], xy[1], rng.uniform(-3, 3)])
    return np.array(pts).ravel()


@pytest.mark.parametrize("name", list(SPECS))
def test_single_triangle_drop_equals_supremum(name):
    spec = SPECS[name]
    cut = O.cutter_from_spec(spec)
    rng = np.random.default_rng(7 + len(name))
    n_checked = n_type = 0
    for _ in range(300):
        tri = random_tri(rng)
        x, y = rng.uniform(-5, 5, 2)
        r = O.drop_triangle(cut, tri, x, y, -1e9)
        zb, kind = drop_sup(spec, tri, (x, y, 0.0))
        if kind is None:
            assert not r["lifted"]
            continue
        assert r["lifted"], (tri, x, y)
        tol = 2e-8 if spec["kind"] == "bull" else 1e-10
        assert r["z"] == pytest.approx(zb, abs=tol), (name, tri.tolist(), x, y, kind, r)
        n_checked += 1
        # where the sup is clearly unique in kind, the reported type agrees
        alt = [drop_sup_kind_excluding(spec, tri, (x, y), k) for k in ("facet", "vertex", "edge") if k != kind]
        if all(a < zb - 1e-6 for a in alt):
            assert TYPE_KIND[r["type"]] == kind, (name, tri.tolist(), x, y, kind, r)
            n_type += 1
    assert n_checked > 100 and n_type > 50


def drop_sup_kind_excluding(spec, tri, cl, kind):
    """Best value among contacts of the given kind only (-inf if none)."""
    from bruteforce import _edge_max, facet_stationary, profile
    t = np.asarray(tri, dtype=float).reshape(3, 3)
    R, h = profile(spec)
    cl3 = np.array([cl[0], cl[1], 0.0])
    vals = [-math.inf]
    if kind == "vertex":
        for v in t:
            q = math.hypot(v[0] - cl[0], v[1] - cl[1])
            if q <= R:
                vals.append(v[2] - h(q))
    elif kind == "edge":
        for i in range(3):
            e = _edge_max(t[i], t[(i + 1) % 3], cl3, R, h)
            if e is not None:
                # an edge maximum at an endpoint is a vertex contact
                p = e[1]
                if min(np.linalg.norm(p - t[i]), np.linalg.norm(p - t[(i + 1) % 3])) > 1e-6:
                    vals.append(e[0])
    else:
        fs = facet_stationary(spec, t, cl3)
        if fs is not None:
            vals.append(fs[0])
    return max(vals)


@pytest.fixture(scope="module")
def c0_small():
    tris = W.heightfield(24, 40.0, bumps=3, steps=2, seed=5)
    cl = W.subsample(W.cl_grid(80, 80, 0.5, z0=-10.0), 1500, seed=3)
    return tris, cl


@pytest.mark.parametrize("name", ["cyl", "ball", "bull", "cone_steep"])
def test_batch_search_modes_agree(c0_small, name):
    """Reference KD search (mode 0), overlap scan in KD rank (mode 1) and all
    triangles (mode 2) give the same heights; modes 0 and 1 also give the
    same contact triangle, type and CC (identical evaluation order).  Cone
    heights depend on the order: a tip or rim facet lift that loses to the
    current z lets the vertex and edge tests run (conecutter.cpp:100-102,
    millingcutter.cpp:386-409), so mode 2 is compared for the other cutters."""
    tris, cl = c0_small
    cut = O.cutter_from_spec(SPECS[name])
    m = O.Mesh(tris)
    r0 = m.batch_drop(cut, cl, search_mode=0)
    r1 = m.batch_drop(cut, cl, search_mode=1)
    r2 = m.batch_drop(cut, cl, search_mode=2)
    assert np.array_equal(r0["z"], r1["z"])
    assert np.array_equal(r0["type"], r1["type"])
    assert np.array_equal(r0["tri"], r1["tri"])
    assert np.array_equal(r0["cc"], r1["cc"])
    if SPECS[name]["kind"] != "cone":
        assert np.array_equal(r0["z"], r2["z"])


def test_batch_drop_brute_force_mesh(c0_small):
    """Whole-mesh drop = max over triangles of the per-triangle supremum."""
    tris, cl = c0_small
    spec = SPECS["bull"]
    r = O.Mesh(tris).batch_drop(O.cutter_from_spec(spec), cl[:60], search_mode=1)
    R = spec["d"] / 2
    for k in range(60):
        x, y, z0 = cl[k]
        near = [t for t in tris if (t[0::3].min() <= x + R and t[0::3].max() >= x - R and
                                    t[1::3].min() <= y + R and t[1::3].max() >= y - R)]
        zb = max([z0] + [drop_sup(spec, t, (x, y, 0.0))[0] for t in near])
        assert r["z"][k] == pytest.approx(zb, abs=2e-8)


def test_kd_rank_is_a_permutation_and_search_is_superset(c0_small):
    tris, cl = c0_small
    m = O.Mesh(tris)
    rank = m.kd_rank(0)
    assert sorted(rank.tolist()) == list(range(len(tris)))
    cut = O.cutter_from_spec(SPECS["ball"])
    R = 3.0
    for k in range(0, 1500, 97):
        x, y, z = cl[k]
        got = m.kd_search_drop(cut, x, y, z)
        # KD order is hi-first DFS: ranks strictly increase along the search
        assert np.all(np.diff(rank[got]) > 0)
        ov = [i for i, t in enumerate(tris) if (t[0::3].min() <= x + R and t[0::3].max() >= x - R and
                                                t[1::3].min() <= y + R and t[1::3].max() >= y - R)]
        assert set(ov) <= set(got.tolist())


# --------------------------------------------------------------- push pins
PUSH_SPECS = {
    "cyl": {"kind": "cyl", "d": 6.0, "L": 20.0},
    "ball": {"kind": "ball", "d": 6.0, "L": 20.0},
    "bull": {"kind": "bull", "d": 6.0, "r": 1.2, "L": 20.0},
    "cone": {"kind": "cone", "d": 6.0, "angle": math.radians(50), "L": 10.0},
}


def collide_interval(spec, tri, fib, fz, samples=801):
    """Interval of fiber parameters where the drop height exceeds fz."""
    p1, p2 = np.asarray(fib[:3]), np.asarray(fib[3:])
    ts = np.linspace(0, 1, samples)
    vals = np.array([drop_sup(spec, tri, p1 + t * (p2 - p1), geometric=True)[0] - fz for t in ts])
    inside = vals > 0
    if not inside.any():
        return None
    idx = np.nonzero(inside)[0]
    i0, i1 = idx[0], idx[-1]

    def g(t):
        return drop_sup(spec, tri, p1 + t * (p2 - p1), geometric=True)[0] - fz

    def bisect(a, b):  # g(a) <= 0 < g(b) or reverse
        ga = g(a)
        for _ in range(60):
            m = 0.5 * (a + b)
            if (g(m) > 0) == (ga > 0):
                a = m
            else:
                b = m
        return 0.5 * (a + b)
    lo = 0.0 if i0 == 0 else bisect(ts[i0 - 1], ts[i0])
    hi = 1.0 if i1 == len(ts) - 1 else bisect(ts[i1], ts[i1 + 1])
    return lo, hi


@pytest.mark.parametrize("name", list(PUSH_SPECS))
def test_push_interval_matches_drop_crossings(name):
    spec = PUSH_SPECS[name]
    cut = O.cutter_from_spec(spec)
    rng = np.random.default_rng(11 + len(name))
    fz = 0.0
    checked = 0
    for it in range(40):
        tri = random_tri(rng)
        tri[2::3] = rng.uniform(-1.0, 6.0, 3)  # within [fz-1, fz+L)
        y = rng.uniform(-4, 4)
        if it % 2 == 0:
            fib = np.array([-15, y, fz, 15, y, fz], dtype=float)
        else:
            fib = np.array([y, -15, fz, y, 15, fz], dtype=float)
        ne, out = O.push_triangle(cut, fib, tri)
        ref = collide_interval(spec, tri, fib, fz)
        if ref is None:
            assert not ne or out[1] - out[0] < 1e-6
            continue
        if ref[1] - ref[0] < 1e-4:
            continue  # grazing contact: the sampled reference cannot resolve it
        assert ne, (name, tri.tolist(), fib.tolist(), ref)
        tol = 1e-6
        assert out[0] == pytest.approx(ref[0], abs=tol) and out[1] == pytest.approx(ref[1], abs=tol), \
            (name, tri.tolist(), fib.tolist(), ref, out[:2])
        checked += 1
    assert checked >= 15


def test_fiber_union_matches_sort_merge():
    rng = np.random.default_rng(3)
    for _ in range(50):
        n = rng.integers(1, 30)
        lo = rng.uniform(0, 1, n)
        hi = lo + rng.uniform(0, 0.2, n)
        if n > 3:  # exact touching pairs merge
            lo[1] = hi[0]
        ints = np.zeros((n, 10))
        ints[:, 0], ints[:, 1] = lo, hi
        ints[:, 5] = ints[:, 9] = 1  # VERTEX: a CC type other than NONE seeds the hull
        got = O.fiber_union(ints)
        ref = union_sorted(zip(lo, hi))
        got_iv = sorted((a, b) for a, b in got[:, :2])
        assert len(got_iv) == len(ref)
        for (a, b), (c, d) in zip(got_iv, ref):
            assert a == c and b == d
