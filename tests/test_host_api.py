"""The drop-in host API (ocl:: C++ classes and their Python binding).

CPU tests cover the host logic: geometry value types, the STL reader, the
cutter constructors, Interval / Fiber::addInterval (checked against the
oracle's union) and paths.  GPU tests run the operations end to end through
the binding and compare them with the oracle and with the C-ABI results.
"""
import math
import os
import struct

import numpy as np
import pytest

from cutters import ALL
from opencamlib_amd import capi, ocl, workloads as W
from oracle import oracle as O

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def surf_from(tris):
    s = ocl.STLSurf()
    for t in np.asarray(tris).reshape(-1, 9):
        s.addTriangle(ocl.Triangle(ocl.Point(*t[0:3]), ocl.Point(*t[3:6]), ocl.Point(*t[6:9])))
    return s


def cutter_from(spec):
    k = spec["kind"]
    if k == "cyl":
        return ocl.CylCutter(spec["d"], spec["L"])
    if k == "ball":
        return ocl.BallCutter(spec["d"], spec["L"])
    if k == "bull":
        return ocl.BullCutter(spec["d"], spec["r"], spec["L"])
    if k == "cone":
        return ocl.ConeCutter(spec["d"], spec["angle"], spec["L"])
    c = ocl.CompositeCutter()
    for s in spec["subs"]:
        c.addCutter(cutter_from(s["cutter"]), s["r"], s["h"], s["zoff"])
    c.setLength(spec["L"])
    return c


# ----------------------------------------------------------------- CPU

def test_point_and_triangle_basics():
    p = ocl.Point(1, 2, 3)
    assert (p * 2).norm() == pytest.approx(2 * math.sqrt(14))
    assert (p - ocl.Point(1, 2, 0)).xyNorm() == 0
    t = ocl.Triangle(ocl.Point(0, 0, 0), ocl.Point(1, 0, 0), ocl.Point(0, 1, 0))
    # (p0-p1) x (p0-p2) = (-1,0,0) x (0,-1,0) = (0,0,1)
    assert (t.n.x, t.n.y, t.n.z) == (0.0, 0.0, 1.0)
    assert ocl.Point(0.5, -1, 0).isRight(ocl.Point(0, 0, 0), ocl.Point(1, 0, 0))
    s = surf_from(W.heightfield(4, 10.0, bumps=1, seed=3))
    assert s.size() == 32
    b = s.getBounds()
    assert b[0] == 0.0 and b[1] == 10.0 and b[2] == 0.0 and b[3] == 10.0


def test_cutter_constructors_and_errors():
    c = ocl.ConeCutter(10, math.pi / 4)  # l defaults to 10: length = r/tan(a) + l
    assert c.getLength() == pytest.approx(15.0)
    assert ocl.BullCutter(10, 2, 20).getRadius() == 5.0
    with pytest.raises(ocl.Error):
        ocl.BullCutter(10, 5, 20)  # no cylindrical part
    with pytest.raises(ocl.Error):
        ocl.CylCutter(-1, 10)
    comp = ocl.CompositeCutter()
    comp.addCutter(ocl.CylCutter(6, 20), 3.0, 0.0, 0.0)
    comp.addCutter(ocl.BallCutter(10, 20), 5.0, 4.0, -1.0)
    assert comp.getDiameter() == 10.0
    with pytest.raises(ocl.Error):
        comp.addCutter(ocl.CompositeCutter(), 6.0, 1.0, 0.0)


def test_composite_presets_are_continuous():
    """Each preset joins its sub-profiles: sub-cutter n's height at its radius
    limit equals sub-cutter n+1's height there (closed forms of the ball,
    bull and cone profiles)."""
    d1, d2, a = 4.0, 10.0, math.radians(30)
    # BallCone: ball r1 meets the cone tangentially at rc = r1 cos(a)
    r1 = d1 / 2
    rc = r1 * math.cos(a)
    h_ball = r1 - math.sqrt(r1 * r1 - rc * rc)
    cone_off = -(rc / math.tan(a) - h_ball)
    assert rc / math.tan(a) + cone_off == pytest.approx(h_ball)
    bc = ocl.BallConeCutter(d1, d2, a)
    assert bc.getRadius() == d2 / 2 and bc.getLength() == 30


def test_interval_first_update_sets_both_ends():
    i = ocl.Interval()
    assert i.empty()
    a = ocl.CCPoint(1, 1, 1)
    a.type = ocl.CCType.VERTEX
    i.updateUpper(2.0, a)
    assert (i.lower, i.upper) == (2.0, 2.0) and i.lower_cc.type == ocl.CCType.VERTEX
    b = ocl.CCPoint(0, 0, 0)
    b.type = ocl.CCType.FACET
    i.updateLower(1.0, b)
    assert (i.lower, i.upper) == (1.0, 2.0)
    assert i.lower_cc.type == ocl.CCType.FACET and i.upper_cc.type == ocl.CCType.VERTEX
    # a CC point of type NONE seeds the interval again (interval.cpp:61-80)
    j = ocl.Interval()
    j.updateUpper(3.0, ocl.CCPoint(1, 1, 1))
    j.updateLower(1.0, ocl.CCPoint(0, 0, 0))
    assert (j.lower, j.upper) == (1.0, 1.0)


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_fiber_add_interval_matches_oracle_union(seed):
    rng = np.random.default_rng(seed)
    n = 40
    lo = rng.uniform(0, 1, n).round(2)  # rounded: equal ends and touching intervals occur
    hi = lo + 0.01 + rng.uniform(0, 0.15, n).round(2)  # lower < upper: both CCs kept
    rows = np.zeros((n, 10))
    rows[:, 0], rows[:, 1] = lo, hi
    rows[:, 2:5] = rng.normal(size=(n, 3))
    rows[:, 5] = rng.integers(1, 16, n)
    rows[:, 6:9] = rng.normal(size=(n, 3))
    rows[:, 9] = rng.integers(1, 16, n)
    f = ocl.Fiber(ocl.Point(0, 0, 1), ocl.Point(1, 0, 1))
    for r in rows:
        lc = ocl.CCPoint(*r[2:5]); lc.type = ocl.CCType(int(r[5]))
        uc = ocl.CCPoint(*r[6:9]); uc.type = ocl.CCType(int(r[9]))
        iv = ocl.Interval()
        iv.updateLower(r[0], lc)
        iv.updateUpper(r[1], uc)
        f.addInterval(iv)
    got = np.array([[i.lower, i.upper, i.lower_cc.x, int(i.lower_cc.type), i.upper_cc.x, int(i.upper_cc.type)]
                    for i in f.getInts()])
    ref = O.fiber_union(rows)
    ref = ref[:, [0, 1, 2, 5, 6, 9]]
    assert np.array_equal(got, ref)


def _write_binary_stl(path, tris):
    with open(path, "wb") as fh:
        fh.write(b"binary header".ljust(80, b" "))
        fh.write(struct.pack("<I", len(tris)))
        for t in tris:
            fh.write(struct.pack("<3f", 0, 0, 1))
            fh.write(struct.pack("<9f", *t))
            fh.write(struct.pack("<H", 0))


def test_stl_reader_binary_and_ascii(tmp_path):
    tris = W.heightfield(3, 7.0, bumps=1, seed=9)
    p = str(tmp_path / "m.stl")
    _write_binary_stl(p, tris)
    s = ocl.STLSurf()
    ocl.STLReader(p, s)
    got = np.array([[c for q in t.getPoints() for c in (q.x, q.y, q.z)] for t in s.getTriangles()])
    assert np.array_equal(got, tris.astype(np.float32).astype(np.float64))
    # text STL: the reference's carpet1.stl data file
    s2 = ocl.STLSurf()
    ocl.STLReader(os.path.join(GOLDEN, "carpet1.stl"), s2)
    verts = []
    with open(os.path.join(GOLDEN, "carpet1.stl")) as fh:
        for line in fh:
            w = line.split()
            if w and w[0] == "vertex":
                verts.append([float(np.float32(v)) for v in w[1:4]])
    verts = np.array(verts).reshape(-1, 9)
    got2 = np.array([[c for q in t.getPoints() for c in (q.x, q.y, q.z)] for t in s2.getTriangles()])
    assert s2.size() == 110
    assert np.array_equal(got2, verts)
    s3 = ocl.STLSurf()
    ocl.STLReader(str(tmp_path / "missing.stl"), s3)
    assert s3.size() == 0


@pytest.mark.parametrize("ccw", [True, False])
def test_arc_points_closed_form(ccw):
    a = ocl.Arc(ocl.Point(2, 0, 1), ocl.Point(0, 2, 1), ocl.Point(0, 0, 1), ccw)
    # counter-clockwise: a quarter circle of radius 2; clockwise: three quarters
    ang = math.pi / 2 if ccw else 3 * math.pi / 2
    assert a.length2d() == pytest.approx(2 * ang)
    for t in (0.0, 0.25, 0.5, 0.9, 1.0):
        p = a.getPoint(t)
        th = (ang if ccw else -ang) * t
        assert (p.x, p.y, p.z) == pytest.approx((2 * math.cos(th), 2 * math.sin(th), 1.0), abs=1e-12)
    ln = ocl.Line(ocl.Point(0, 0, 0), ocl.Point(3, 4, 5))
    assert ln.length2d() == 5.0
    pth = ocl.Path()
    pth.append(a)
    pth.append(ln)
    assert len(ocl.Path(pth).getSpans()) == 2


# ----------------------------------------------------------------- GPU

@pytest.fixture(scope="module")
def scene():
    return W.heightfield(24, 30.0, bumps=3, steps=2, seed=21)


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["cyl", "ball", "bull", "cone", "cylball"])
def test_batch_drop_cutter_class_matches_oracle(gpu, scene, name):
    spec = dict(ALL[name])
    cl = W.cl_grid(30, 30, 1.0, z0=-10.0)
    s = surf_from(scene)
    bdc = ocl.BatchDropCutter()
    bdc.setSTL(s)
    bdc.setCutter(cutter_from(spec))
    bdc.setPrecision(capi.PREC["exact64"])
    for p in cl:
        bdc.appendPoint(ocl.CLPoint(*p))
    bdc.run()
    pts = bdc.getCLPoints()
    z = np.array([p.z for p in pts])
    ty = np.array([int(p.cc().type) for p in pts])
    ref = O.Mesh(scene).batch_drop(O.cutter_from_spec(spec), cl, search_mode=1)
    tol = 1e-12 if name == "bull" else 0.0
    assert np.max(np.abs(z - ref["z"])) <= tol
    assert np.array_equal(ty, ref["type"])
    assert bdc.getCalls() > 0
    # default precision (mixed): types exact, z within 1e-6 mm
    bdc2 = ocl.BatchDropCutter()
    bdc2.setSTL(s)
    bdc2.setCutter(cutter_from(spec))
    for p in cl:
        bdc2.appendPoint(ocl.CLPoint(*p))
    bdc2.run()
    pts2 = bdc2.getCLPoints()
    assert np.max(np.abs(np.array([p.z for p in pts2]) - ref["z"])) <= 1e-6
    assert np.array_equal(np.array([int(p.cc().type) for p in pts2]), ref["type"])


@pytest.mark.gpu
def test_point_drop_cutter_and_triangles_under_cutter(gpu, scene):
    spec = ALL["ball"]
    s = surf_from(scene)
    pdc = ocl.PointDropCutter()
    pdc.setSTL(s)
    pdc.setCutter(cutter_from(spec))
    pdc.setPrecision(capi.PREC["exact64"])
    m = O.Mesh(scene)
    cut = O.cutter_from_spec(spec)
    for x, y in [(3.3, 4.1), (15.0, 15.0), (29.0, 0.5)]:
        p = ocl.CLPoint(x, y, -10.0)
        pdc.run(p)
        ref = m.batch_drop(cut, np.array([[x, y, -10.0]]), search_mode=1)
        assert p.z == ref["z"][0] and int(p.cc().type) == ref["type"][0]
    bdc = ocl.BatchDropCutter()
    bdc.setSTL(s)
    for x, y in [(3.3, 4.1), (15.0, 15.0)]:
        got = bdc.getTrianglesUnderCutter(ocl.CLPoint(x, y, -10.0), cutter_from(spec))
        idx = m.kd_search_drop(cut, x, y, -10.0)
        want = [scene[i] for i in idx]
        have = [[c for q in t.getPoints() for c in (q.x, q.y, q.z)] for t in got]
        assert np.array_equal(np.array(have), np.array(want))


@pytest.mark.gpu
@pytest.mark.parametrize("direction", ["x", "y"])
def test_batch_push_cutter_class_matches_capi(gpu, direction):
    tris = W.spheres_scene(3, 2, 50.0, seed=4)
    spec = ALL["bull"]
    fib = np.concatenate([W.waterline_fibers(tris, 5.0, z, 40)[0 if direction == "x" else 1]
                          for z in (2.0, 7.0)])
    bpc = ocl.BatchPushCutter()
    (bpc.setXDirection if direction == "x" else bpc.setYDirection)()
    bpc.setSTL(surf_from(tris))
    bpc.setCutter(cutter_from(spec))
    for f in fib:
        bpc.appendFiber(ocl.Fiber(ocl.Point(*f[:3]), ocl.Point(*f[3:])))
    bpc.run()
    fibers = bpc.getFibers()
    dev = capi.DeviceMesh(tris)
    ref = dev.push(spec, fib, direction)
    dev.close()
    got = np.array([[i.lower, i.upper] for f in fibers for i in f.getInts()]).reshape(-1, 2)
    assert np.array_equal(got, ref["lohi"])
    assert len(bpc.getCLPoints()) == 2 * len(got)
    # the overlap query is the reference's KD search for the fiber
    m = O.Mesh(tris)
    f0 = fib[len(fib) // 2]
    tri_list = bpc.getOverlapTriangles(ocl.Fiber(ocl.Point(*f0[:3]), ocl.Point(*f0[3:])))
    idx = m.kd_search_push(O.cutter_from_spec(spec), f0, 0 if direction == "x" else 1)
    have = np.array([[c for q in t.getPoints() for c in (q.x, q.y, q.z)] for t in tri_list]).reshape(-1, 9)
    assert np.array_equal(have, tris[idx].reshape(-1, 9))


def _span_points_line(p1, p2, sampling):
    length = math.hypot(p2[0] - p1[0], p2[1] - p1[1])
    steps = int(length / sampling + 1)
    return [[p1[k] + (p2[k] - p1[k]) * (i / steps) for k in range(3)] for i in range(steps + 1)]


@pytest.mark.gpu
def test_path_drop_cutter_matches_oracle(gpu, scene):
    spec = ALL["cyl"]
    lines = [((1, 1, 0), (28, 3, 0)), ((28, 3, 0), (5, 25, 0))]
    pth = ocl.Path()
    for a, b in lines:
        pth.append(ocl.Line(ocl.Point(*a), ocl.Point(*b)))
    pdc = ocl.PathDropCutter()
    pdc.setSTL(surf_from(scene))
    pdc.setCutter(cutter_from(spec))
    pdc.setPath(pth)
    pdc.setSampling(0.3)
    pdc.setZ(-5.0)
    pdc.setPrecision(capi.PREC["exact64"])
    pdc.run()
    pts = pdc.getCLPoints()
    want = np.array([q for a, b in lines for q in _span_points_line(a, b, 0.3)])
    want[:, 2] = -5.0
    got = np.array([[p.x, p.y] for p in pts])
    assert np.array_equal(got, want[:, :2])
    ref = O.Mesh(scene).batch_drop(O.cutter_from_spec(spec), want, search_mode=1)
    assert np.array_equal(np.array([p.z for p in pts]), ref["z"])


@pytest.mark.gpu
def test_adaptive_path_drop_cutter_matches_recursive_oracle(gpu, scene):
    """Level-by-level device batches give the reference's depth-first point
    list (adaptivepathdropcutter.cpp:72-94), rebuilt here recursively with
    single-point oracle drops."""
    spec = ALL["ball"]
    a, b = (2.0, 3.0, -2.0), (27.0, 22.0, -2.0)
    pth = ocl.Path()
    pth.append(ocl.Line(ocl.Point(*a), ocl.Point(*b)))
    apdc = ocl.AdaptivePathDropCutter()
    apdc.setSTL(surf_from(scene))
    apdc.setCutter(cutter_from(spec))
    apdc.setPath(pth)
    apdc.setSampling(1.0)
    apdc.setMinSampling(0.05)
    apdc.setCosLimit(0.999)
    apdc.setPrecision(capi.PREC["exact64"])
    apdc.run()
    got = np.array([[p.x, p.y, p.z] for p in apdc.getCLPoints()])

    m = O.Mesh(scene)
    cut = O.cutter_from_spec(spec)

    def at(t):
        q = [a[k] + (b[k] - a[k]) * t for k in range(3)]
        r = m.batch_drop(cut, np.array([q]), search_mode=1)
        return np.array([q[0], q[1], r["z"][0]])

    def flat(p0, p1, p2):
        # Point::normalize scales by the reciprocal of the 3D norm
        v1, v2 = p1 - p0, p2 - p1
        n1 = math.sqrt(v1[0] * v1[0] + v1[1] * v1[1] + v1[2] * v1[2])
        n2 = math.sqrt(v2[0] * v2[0] + v2[1] * v2[1] + v2[2] * v2[2])
        v1 = v1 * (1 / n1) if n1 else v1
        v2 = v2 * (1 / n2) if n2 else v2
        return v1[0] * v2[0] + v1[1] * v2[1] + v1[2] * v2[2] > 0.999

    out = []

    def rec(t0, t1, c0, c1):
        tm = t0 + (t1 - t0) / 2.0
        cm = at(tm)
        d = c1 - c0
        step = math.sqrt(d[0] * d[0] + d[1] * d[1])
        if step > 1.0 or (not flat(c0, cm, c1) and step > 0.05):
            rec(t0, tm, c0, cm)
            rec(tm, t1, cm, c1)
        else:
            out.append(c1)

    c0, c1 = at(0.0), at(1.0)
    out.append(c0)
    rec(0.0, 1.0, c0, c1)
    want = np.array(out)
    assert got.shape == want.shape
    assert np.array_equal(got, want)


random_polylines = W.random_polylines


def refine_and_measure(tris, spec, polylines, tol, sampling, min_sampling, z0=-20.0):
    """Slope-driven refinement of each polyline, then the true chord error of
    every output interval from 15 extra drops inside it."""
    s = surf_from(tris)
    cut = cutter_from(spec)
    pts_all, segs = [], []
    levels = 0
    for poly in polylines:
        pth = ocl.Path()
        for a, b in zip(poly[:-1], poly[1:]):
            pth.append(ocl.Line(ocl.Point(a[0], a[1], z0), ocl.Point(b[0], b[1], z0)))
        apdc = ocl.AdaptivePathDropCutter()
        apdc.setSTL(s)
        apdc.setCutter(cut)
        apdc.setPath(pth)
        apdc.setSampling(sampling)
        apdc.setMinSampling(min_sampling)
        apdc.setChordTolerance(tol)
        apdc.run()
        levels = max(levels, apdc.getLevels())
        p = np.array([[q.x, q.y, q.z] for q in apdc.getCLPoints()])
        pts_all.append(p)
        for i in range(len(p) - 1):
            if np.hypot(*(p[i + 1, :2] - p[i, :2])) > 0:
                segs.append((p[i], p[i + 1]))
    a = np.array([s_[0] for s_ in segs])
    b = np.array([s_[1] for s_ in segs])
    tau = np.arange(1, 16) / 16.0
    probe = (a[:, None, :2] + (b - a)[:, None, :2] * tau[None, :, None]).reshape(-1, 2)
    cl = np.concatenate([probe, np.full((len(probe), 1), z0)], axis=1)
    dev = capi.DeviceMesh(tris)
    zt = dev.drop(spec, cl, "mixed")["z"].reshape(len(segs), 15)
    dev.close()
    chord = a[:, None, 2] + (b - a)[:, None, 2] * tau[None, :]
    err = np.max(np.abs(zt - chord), axis=1)
    length = np.hypot(*(b - a)[:, :2].T)
    return sum(len(p) for p in pts_all), err, length, levels


@pytest.mark.gpu
def test_slope_driven_refinement_meets_chord_tolerance(gpu, scene):
    """setChordTolerance replaces the flatness test: intervals longer than
    the minimum step meet the chord tolerance (checked by dense drops), and
    the refinement uses far fewer points than uniform sampling at the
    minimum step."""
    tol, min_step = 0.01, 0.01
    polys = random_polylines(20, 30.0, seed=7)
    npts, err, length, levels = refine_and_measure(scene, ALL["ball"], polys, tol, 1.0, min_step)
    long_ = length > 2 * min_step
    frac = float(np.mean(err[long_] <= tol))
    total_len = sum(np.sum(np.hypot(*np.diff(p, axis=0).T)) for p in polys)
    print(f"[refine] points={npts} intervals={len(err)} levels={levels} max_err={err.max():.4f} "
          f"p99_err={np.percentile(err, 99):.4f} frac_ok(long)={frac:.4f} "
          f"uniform_points_at_min_step={int(total_len / min_step)}", flush=True)
    assert frac >= 0.995 and err[long_].max() <= 1.5 * tol
    assert npts < total_len / min_step / 5


def test_waterline_loops_are_not_built():
    wl = ocl.Waterline()
    with pytest.raises(ocl.Error, match="weave"):
        wl.getLoops()


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["cone", "bull"])
def test_waterline_fibers_match_oracle(gpu, name):
    """Waterline::run (both directions in one launch) against the oracle's
    batch push of the same fibers, per direction: counts, endpoints (Bull
    within 1e-9) and CC types."""
    tris = W.spheres_scene(3, 2, 50.0, seed=4)
    spec = ALL[name]
    s = surf_from(tris)
    wl = ocl.Waterline()
    wl.setSTL(s)
    wl.setCutter(cutter_from(spec))
    wl.setZ(3.5)
    wl.setSampling(0.7)
    wl.run()
    om = O.Mesh(tris)
    tol = 1e-9 if name == "bull" else 0.0
    for fibers, d in ((wl.getXFibers(), 0), (wl.getYFibers(), 1)):
        arr = np.array([[f.p1.x, f.p1.y, f.p1.z, f.p2.x, f.p2.y, f.p2.z] for f in fibers])
        ref = om.batch_push(O.cutter_from_spec(spec), arr, d, search_mode=0)
        offs = ref["offsets"]
        nint = 0
        for i, f in enumerate(fibers):
            got = sorted((iv.lower, iv.upper, int(iv.lower_cc.type), int(iv.upper_cc.type)) for iv in f.getInts())
            want = sorted((r[0], r[1], int(c[3]), int(c[7])) for r, c in
                          zip(ref["lohi"][offs[i]:offs[i + 1]], ref["cc"][offs[i]:offs[i + 1]]))
            assert len(got) == len(want), (d, i)
            for g, w in zip(got, want):
                assert abs(g[0] - w[0]) <= tol and abs(g[1] - w[1]) <= tol and g[2:] == w[2:], (d, i, g, w)
            nint += len(got)
        assert nint > 10


@pytest.mark.gpu
def test_push_xy_equals_two_single_direction_pushes(gpu):
    """ocl_hip_push_xy's rows per direction equal ocl_hip_push_batch's for
    that direction alone, bit for bit."""
    tris = W.spheres_scene(4, 2, 60.0, seed=5)
    spec = ALL["bull"]
    xf, yf = W.waterline_fibers(tris, 5.0, 4.0, 80)
    dev = capi.DeviceMesh(tris)
    both = dev.push_xy(spec, xf, yf)
    for d, fib in (("x", xf), ("y", yf)):
        one = dev.push(spec, fib, d)
        for k in ("offsets", "lohi", "cc"):
            assert np.array_equal(both[d][k], one[k]), (d, k)
    assert both["stats"]["points"] == len(xf) + len(yf)
    empty = dev.push_xy(spec, xf, np.zeros((0, 6)))
    assert np.array_equal(empty["x"]["lohi"], both["x"]["lohi"]) and len(empty["y"]["lohi"]) == 0
    dev.close()


@pytest.mark.gpu
def test_waterline_fibers_match_push(gpu):
    """Waterline::init_fibers' grid (bbox +- 2R, `sampling` apart) pushed in
    both directions equals the C-ABI push of the same fibers."""
    tris = W.spheres_scene(3, 2, 50.0, seed=4)
    spec = ALL["cone"]
    s = surf_from(tris)
    wl = ocl.Waterline()
    wl.setSTL(s)
    wl.setCutter(cutter_from(spec))
    wl.setZ(3.5)
    wl.setSampling(0.7)
    wl.run()
    xf = wl.getXFibers()
    yf = wl.getYFibers()
    r = spec["d"] / 2
    b = s.getBounds()
    nx, ny = int((b[1] + 2 * r - (b[0] - 2 * r)) / 0.7), int((b[3] + 2 * r - (b[2] - 2 * r)) / 0.7)
    assert len(xf) == ny + 1 and len(yf) == nx + 1
    dev = capi.DeviceMesh(tris)
    for fibers, d in ((xf, "x"), (yf, "y")):
        arr = np.array([[f.p1.x, f.p1.y, f.p1.z, f.p2.x, f.p2.y, f.p2.z] for f in fibers])
        ref = dev.push(spec, arr, d)
        got = np.array([[i.lower, i.upper] for f in fibers for i in f.getInts()]).reshape(-1, 2)
        assert np.array_equal(got, ref["lohi"])
        assert len(got) > 10
    dev.close()
