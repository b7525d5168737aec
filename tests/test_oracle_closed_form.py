"""Pins of the FP64 oracle by closed forms (no reference fixture exists for
the drop/push path: the reference's own tests assert nothing).

Facet (millingcutter.cpp:51-75, conecutter.cpp:70-105), vertex
(millingcutter.cpp:36-47) and Cyl/Ball edge (cylcutter.cpp:52-70,
ballcutter.cpp:50-63) contacts against analytic solutions of the contact
geometry, plus the two known-answer values measured on the compiled
reference (CylCutter / BallCutter d=0.1 on the plane x+y+z=1 at (0.2, 0.2)).
"""
import math

import numpy as np
import pytest

from oracle import oracle as O

FACET, FACET_TIP, VERTEX, EDGE = 14, 15, 1, 3


def big_plane_tri(gx, gy, c, half=50.0):
    """Triangle of the plane z = c + gx x + gy y covering [-half/2, half/2]^2."""
    pts = [(-half, -half), (2 * half, -half), (-half, 2 * half)]
    return np.array([[x, y, c + gx * x + gy * y] for x, y in pts]).ravel()


def drop(spec, tri, x, y):
    return O.drop_triangle(O.cutter_from_spec(spec), tri, x, y, -1e9)


def test_known_answers_plane_xyz1():
    tri = np.array([-5, -5, 11, 10, -5, -4, -5, 10, -4], dtype=float)  # x + y + z = 1
    r = drop({"kind": "cyl", "d": 0.1, "L": 1.0}, tri, 0.2, 0.2)
    assert r["type"] == FACET
    assert r["z"] == pytest.approx(0.670710678118655, abs=1e-15)
    assert r["z"] == pytest.approx(1 - 0.4 + 0.05 * math.sqrt(2), abs=1e-15)
    r = drop({"kind": "ball", "d": 0.1, "L": 1.0}, tri, 0.2, 0.2)
    assert r["type"] == FACET
    assert r["z"] == pytest.approx(0.636602540378444, abs=1e-15)


@pytest.mark.parametrize("seed", range(20))
def test_facet_contact_closed_form(seed):
    rng = np.random.default_rng(seed)
    gx, gy = rng.uniform(-1.5, 1.5, 2)
    c = rng.uniform(-5, 5)
    x, y = rng.uniform(-3, 3, 2)
    tri = big_plane_tri(gx, gy, c)
    G = math.hypot(gx, gy)
    pz = c + gx * x + gy * y
    R = 3.0
    # Cyl: highest rim point, z = plane(cl) + R |g|
    r = drop({"kind": "cyl", "d": 2 * R, "L": 10}, tri, x, y)
    assert r["type"] == FACET and r["z"] == pytest.approx(pz + R * G, abs=1e-12)
    # Ball: sphere tangent to the plane, z = plane(cl) + R (sqrt(1+|g|^2) - 1)
    r = drop({"kind": "ball", "d": 2 * R, "L": 10}, tri, x, y)
    assert r["type"] == FACET and r["z"] == pytest.approx(pz + R * (math.sqrt(1 + G * G) - 1), abs=1e-12)
    cc = np.array(r["cc"])
    assert cc[2] == pytest.approx(c + gx * cc[0] + gy * cc[1], abs=1e-12)
    assert math.hypot(cc[0] - x, cc[1] - y) == pytest.approx(R * G / math.sqrt(1 + G * G), abs=1e-12)
    # Bull (torus R, corner rb): z = plane(cl) + (R-rb)|g| + rb (sqrt(1+|g|^2) - 1)
    rb = 1.0
    r = drop({"kind": "bull", "d": 2 * R, "r": rb, "L": 10}, tri, x, y)
    assert r["type"] == FACET
    assert r["z"] == pytest.approx(pz + (R - rb) * G + rb * (math.sqrt(1 + G * G) - 1), abs=1e-12)
    # Cone, half-angle a: tip contact while |g| < cot a
    a = math.radians(30.0)
    r = drop({"kind": "cone", "d": 2 * R, "angle": a, "L": 10}, tri, x, y)
    if G < 1 / math.tan(a) - 1e-6:
        assert r["type"] == FACET_TIP and r["z"] == pytest.approx(pz, abs=1e-12)


def test_horizontal_and_vertical_facets():
    tri = big_plane_tri(0.0, 0.0, 2.5)
    for spec in ({"kind": "cyl", "d": 4, "L": 5}, {"kind": "ball", "d": 4, "L": 5},
                 {"kind": "bull", "d": 4, "r": 1, "L": 5}):
        r = drop(spec, tri, 0.3, -0.2)
        assert r["type"] == FACET and r["z"] == 2.5
        assert tuple(r["cc"]) == (0.3, -0.2, 2.5)
    # vertical triangle: no facet contact; the top edge/vertex decides
    vt = np.array([0, 0, 0, 4, 0, 0, 2, 0, 3], dtype=float)
    r = drop({"kind": "cyl", "d": 2, "L": 5}, vt, 2.0, 0.5)
    assert r["type"] in (VERTEX, EDGE) and r["z"] == pytest.approx(3.0, abs=1e-12)


@pytest.mark.parametrize("kind", ["cyl", "ball", "bull", "cone"])
def test_vertex_contact_closed_form(kind):
    # vertical triangle in the plane y=0 with its apex p0 = (0,0,0) as the
    # highest point; a CL beyond p0 in -x sees only that vertex.
    tri = np.array([0, 0, 0, 3, 0, -0.5, 0, 0, -6], dtype=float)
    spec = {"cyl": {"kind": "cyl", "d": 4, "L": 5}, "ball": {"kind": "ball", "d": 4, "L": 5},
            "bull": {"kind": "bull", "d": 4, "r": 0.5, "L": 5},
            "cone": {"kind": "cone", "d": 4, "angle": math.radians(40), "L": 5}}[kind]
    x, y = -0.9, 0.6
    q = math.hypot(x, y)
    R = 2.0
    h = {"cyl": 0.0, "ball": R - math.sqrt(R * R - q * q),
         "bull": (lambda u: 0.0 if u <= 0 else 0.5 - math.sqrt(0.25 - u * u))(q - 1.5),
         "cone": q / math.tan(math.radians(40))}[kind]
    r = drop(spec, tri, x, y)
    assert r["type"] == VERTEX
    assert r["z"] == pytest.approx(-h, abs=1e-13)
    assert tuple(r["cc"]) == (0.0, 0.0, 0.0)


def ridge(m, half=10.0, below=50.0):
    """Triangle whose edge (p0,p1) runs along y=0 with slope m in x, the third
    vertex far below at -y: for a CL at y>0 the edge is the contact."""
    return np.array([-half, 0, -half * m, half, 0, half * m, 0, -20, -below], dtype=float)


@pytest.mark.parametrize("m", [0.0, 0.3, -0.7, 1.8])
def test_edge_contact_cyl_ball_closed_form(m):
    tri = ridge(m)
    R = 2.5
    for d_ in (0.0, 0.4, 1.7, 2.4):
        x = 0.37
        # Cyl: the highest point of the line inside the disk chord, half-length sqrt(R^2-d^2)
        r = drop({"kind": "cyl", "d": 2 * R, "L": 5}, tri, x, d_)
        expect = m * x + abs(m) * math.sqrt(R * R - d_ * d_)
        assert r["type"] == EDGE, (m, d_)
        assert r["z"] == pytest.approx(expect, abs=1e-12)
        # Ball: sphere tangent to the 3-D line: centre height z_foot + sqrt(1+m^2) sqrt(R^2-d^2)
        r = drop({"kind": "ball", "d": 2 * R, "L": 5}, tri, x, d_)
        expect = m * x + math.sqrt(1 + m * m) * math.sqrt(R * R - d_ * d_) - R
        assert r["type"] == EDGE, (m, d_)
        assert r["z"] == pytest.approx(expect, abs=1e-12)
