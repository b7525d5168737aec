"""Independent geometry used to pin the oracle (and the HIP path).

Nothing here follows the reference's formulas.  The drop height of a convex
cutter with profile height h(q) (q = XY distance from the cutter axis, q <= R)
on a triangle T is

    z(cl) = sup { p.z - h(|p - cl|_xy) : p in T, |p - cl|_xy <= R },

a concave maximisation over a convex set.  Its maximum lies at a vertex, on
an edge (1-D concave search), or at an interior stationary point of the plane
(closed form per cutter below).  Profiles:

    Cyl   h = 0
    Ball  h = R - sqrt(R^2 - q^2)
    Bull  h = 0 for q <= R-r, else r - sqrt(r^2 - (q-R+r)^2)
    Cone  h = q / tan(angle)
    Composite  h = h_n(q) + zoffset_n on the n-th radius window
               (R_{n-1}, R_n] (R_{-1} = 0), R = the last R_n

For a composite the interior maximiser over a plane of uphill slope G sits at
the distance q maximising G q - h(q), a concave function when h is convex:
found by golden-section search, not by any closed form.
"""
from __future__ import annotations

import math

import numpy as np


def _golden_max(f, lo, hi, iters=200):
    """argmax of a concave f on [lo, hi] by golden-section search."""
    g = (math.sqrt(5.0) - 1.0) / 2.0
    x1 = hi - g * (hi - lo)
    x2 = lo + g * (hi - lo)
    f1, f2 = f(x1), f(x2)
    for _ in range(iters):
        if hi - lo < 1e-15:
            break
        if f1 < f2:
            lo, x1, f1 = x1, x2, f2
            x2 = lo + g * (hi - lo)
            f2 = f(x2)
        else:
            hi, x2, f2 = x2, x1, f1
            x1 = hi - g * (hi - lo)
            f1 = f(x1)
    return max((lo, hi, 0.5 * (lo + hi)), key=f)


def profile(spec):
    k = spec["kind"]
    if k == "composite":
        parts = []
        for s in spec["subs"]:
            _, hs = profile(s["cutter"])
            parts.append((s["r"], s["zoff"], hs))
        R = parts[-1][0]

        def hc(q):
            for rn, zo, hs in parts:
                if q <= rn:
                    return hs(q) + zo
            rn, zo, hs = parts[-1]
            return hs(q) + zo
        return R, hc
    R = spec["d"] / 2.0
    if k == "cyl":
        return R, lambda q: 0.0
    if k == "ball":
        return R, lambda q: R - math.sqrt(max(R * R - q * q, 0.0))
    if k == "bull":
        r = spec["r"]
        r1 = R - r

        def h(q):
            if q <= r1:
                return 0.0
            u = q - r1
            return r - math.sqrt(max(r * r - u * u, 0.0))
        return R, h
    if k == "cone":
        cot = 1.0 / math.tan(spec["angle"])
        return R, lambda q: q * cot
    raise ValueError(k)


def _edge_max(a, b, cl, R, h, iters=200):
    """max over t in [0,1] of (a + t(b-a)).z - h(q(t)) restricted to q <= R;
    f is concave on that chord, so golden-section search converges to it."""
    d = b[:2] - a[:2]
    w = a[:2] - cl[:2]
    A = float(d @ d)
    B = 2.0 * float(w @ d)
    Cc = float(w @ w) - R * R
    if A < 1e-30:
        return None
    disc = B * B - 4 * A * Cc
    if disc < 0:
        return None
    s = math.sqrt(disc)
    t0 = max(0.0, (-B - s) / (2 * A))
    t1 = min(1.0, (-B + s) / (2 * A))
    if t0 > t1:
        return None

    def f(t):
        p = a + t * (b - a)
        q = math.hypot(p[0] - cl[0], p[1] - cl[1])
        return p[2] - h(min(q, R)), p

    g = (math.sqrt(5.0) - 1.0) / 2.0
    lo, hi = t0, t1
    x1 = hi - g * (hi - lo)
    x2 = lo + g * (hi - lo)
    f1, f2 = f(x1)[0], f(x2)[0]
    for _ in range(iters):
        if hi - lo < 1e-15:
            break
        if f1 < f2:
            lo, x1, f1 = x1, x2, f2
            x2 = lo + g * (hi - lo)
            f2 = f(x2)[0]
        else:
            hi, x2, f2 = x2, x1, f1
            x1 = hi - g * (hi - lo)
            f1 = f(x1)[0]
    cands = [f(t0), f(t1), f(0.5 * (lo + hi))]
    return max(cands, key=lambda c: c[0])


def _inside(p, tri):
    a, b, c = tri
    def cross(o, u, v):
        return (u[0] - o[0]) * (v[1] - o[1]) - (u[1] - o[1]) * (v[0] - o[0])
    s1, s2, s3 = cross(a, b, p), cross(b, c, p), cross(c, a, p)
    return (s1 > 0 and s2 > 0 and s3 > 0) or (s1 < 0 and s2 < 0 and s3 < 0)


def facet_stationary(spec, tri, cl):
    """Interior maximiser of p.z - h(q) over the (infinite) plane of `tri`,
    or None (vertical plane / cone rim case).  Returns (z, p)."""
    a, b, c = tri
    n = np.cross(b - a, c - a)
    if abs(n[2]) < 1e-12:
        return None
    gx, gy = -n[0] / n[2], -n[1] / n[2]          # plane z = a.z + gx(x-a.x) + gy(y-a.y)
    G = math.hypot(gx, gy)
    plane = lambda x, y: a[2] + gx * (x - a[0]) + gy * (y - a[1])
    ux, uy = (gx / G, gy / G) if G > 0 else (0.0, 0.0)
    R = profile(spec)[0]
    k = spec["kind"]
    if k == "cyl":
        q = R if G > 0 else 0.0
        dz = 0.0
    elif k == "ball":
        s = G / math.sqrt(1 + G * G)
        q = R * s
        dz = R - math.sqrt(R * R - q * q)
    elif k == "bull":
        r = spec["r"]
        s = G / math.sqrt(1 + G * G)
        q = (R - r) * (1 if G > 0 else 0) + r * s
        dz = r - math.sqrt(r * r - (r * s) ** 2)
    elif k == "cone":
        cot = 1.0 / math.tan(spec["angle"])
        if G >= cot:
            q, dz = R, R * cot  # rim contact
        else:
            q, dz = 0.0, 0.0    # tip contact
    elif k == "composite":
        _, h = profile(spec)
        q = _golden_max(lambda u: G * u - h(u), 0.0, R) if G > 0 else 0.0
        dz = h(q)
    else:
        raise ValueError(k)
    px, py = cl[0] + q * ux, cl[1] + q * uy
    pz = plane(px, py)
    p = np.array([px, py, pz])
    if not _inside(p, tri):
        return None
    return pz - dz, p


def cone_facet_rule(spec, tri, cl):
    """The reference's ConeCutter facet drop accepts the tip contact whenever
    the CL's XY lies strictly inside the triangle, and otherwise the rim
    contact whenever the rim point uphill of the CL lies inside, whatever the
    facet slope (conecutter.cpp:70-105); a facet lift then skips the vertex
    and edge tests (millingcutter.cpp:386-409).  Geometric values, with the
    cutter built with l=0 so that `length` equals the cone height."""
    a, b, c = tri
    n = np.cross(b - a, c - a)
    if abs(n[2]) < 1e-7 * np.linalg.norm(n):
        return None
    gx, gy = -n[0] / n[2], -n[1] / n[2]
    G = math.hypot(gx, gy)
    plane = lambda x, y: a[2] + gx * (x - a[0]) + gy * (y - a[1])
    if _inside(cl, tri):
        return plane(cl[0], cl[1]), "facet"
    if G == 0:
        return None
    R = spec["d"] / 2.0
    ch = R / math.tan(spec["angle"])
    px, py = cl[0] + R * gx / G, cl[1] + R * gy / G
    if _inside((px, py), tri):
        return plane(px, py) - ch, "facet"
    return None


def drop_sup(spec, tri9, cl, geometric=False):
    """(z, kind) of the supremum over one triangle; kind in {"facet",
    "vertex", "edge"} names where the maximum sits; z = -inf if the cutter
    does not reach the triangle.  For cones the reference's facet-first rule
    (cone_facet_rule) is applied before the supremum unless `geometric`."""
    tri = np.asarray(tri9, dtype=np.float64).reshape(3, 3)
    cl = np.asarray(cl, dtype=np.float64)
    R, h = profile(spec)
    if spec["kind"] == "cone" and not geometric:
        fr = cone_facet_rule(spec, tri, cl)
        if fr is not None:
            return fr
    best = [(-math.inf, None)]
    for v in tri:
        q = math.hypot(v[0] - cl[0], v[1] - cl[1])
        if q <= R:
            best.append((v[2] - h(q), "vertex"))
    for i in range(3):
        e = _edge_max(tri[i], tri[(i + 1) % 3], cl, R, h)
        if e is not None:
            best.append((e[0], "edge"))
    if spec["kind"] != "cone" or geometric:
        fs = facet_stationary(spec, tri, cl)
        if fs is not None:
            best.append((fs[0], "facet"))
    return max(best, key=lambda x: x[0])


def union_sorted(intervals):
    """Union of closed intervals [lo, hi] by sort and sweep; touching
    intervals merge.  Returns a sorted list of (lo, hi)."""
    iv = sorted((lo, hi) for lo, hi in intervals)
    out = []
    for lo, hi in iv:
        if out and lo <= out[-1][1]:
            out[-1] = (out[-1][0], max(out[-1][1], hi))
        else:
            out.append((lo, hi))
    return out
