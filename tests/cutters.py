"""Cutter specs used by tests and the bench (shared dictionaries, see
opencamlib_amd.capi.cutter_desc)."""
import math

CYL6 = {"kind": "cyl", "d": 6.0, "L": 20.0}
BALL10 = {"kind": "ball", "d": 10.0, "L": 20.0}
BULL10 = {"kind": "bull", "d": 10.0, "r": 2.0, "L": 20.0}
CONE10 = {"kind": "cone", "d": 10.0, "angle": math.pi / 4, "L": 10.0}
# Cyl core d=6 + Ball rim d=10, the ball shifted down so the profile is
# continuous at r=3: ball height at r=3 is 5 - sqrt(25-9) = 1.
CYLBALL = {"kind": "composite", "L": 20.0, "subs": [
    {"cutter": {"kind": "cyl", "d": 6.0, "L": 20.0}, "r": 3.0, "h": 0.0, "zoff": 0.0},
    {"cutter": {"kind": "ball", "d": 10.0, "L": 20.0}, "r": 5.0, "h": 4.0, "zoff": -1.0},
]}

ALL = {"cyl": CYL6, "ball": BALL10, "bull": BULL10, "cone": CONE10, "cylball": CYLBALL}


def scaled(spec, d):
    """Same family with diameter d (composites keep their fixed geometry)."""
    if spec["kind"] == "composite":
        return spec
    s = dict(spec)
    if s["kind"] == "bull":
        s["r"] = min(s["r"], 0.4 * d)
    s["d"] = d
    return s


def _cone(d, a, l=10.0):
    return {"kind": "cone", "d": d, "angle": a, "L": l}


def cyl_cone(d1, d2, a, l=10.0):
    """CylConeCutter's sub-cutters (compositecutter.cpp:310-325): the cone
    sits below the cylinder's plane (negative offset).  l = the cone
    sub-cutter's own length argument (the reference's default is 10)."""
    off = -(d1 / 2.0) / math.tan(a)
    ch = (d2 / 2.0) / math.tan(a) + off
    return {"kind": "composite", "L": ch + 10.0, "subs": [
        {"cutter": {"kind": "cyl", "d": d1, "L": 1.0}, "r": d1 / 2.0, "h": 0.0, "zoff": 0.0},
        {"cutter": _cone(d2, a, l), "r": d2 / 2.0, "h": ch, "zoff": off},
        {"cutter": {"kind": "cyl", "d": d2, "L": 20.0}, "r": d2 / 2.0, "h": (d2 / 2.0) / math.tan(a) + 20.0,
         "zoff": ch}]}


def ball_cone(d1, d2, a, l=10.0):
    """BallConeCutter's sub-cutters (compositecutter.cpp:328-347): the ball
    meets the cone tangentially at r1 cos(a)."""
    r1, r2 = d1 / 2.0, d2 / 2.0
    rc = r1 * math.cos(a)
    h1 = r1 - math.sqrt(r1 * r1 - rc * rc)
    off = -(rc / math.tan(a) - h1)
    h2 = r2 / math.tan(a) + off
    return {"kind": "composite", "L": 30.0, "subs": [
        {"cutter": {"kind": "ball", "d": d1, "L": 1.0}, "r": rc, "h": h1, "zoff": 0.0},
        {"cutter": _cone(d2, a, l), "r": r2, "h": h2, "zoff": off},
        {"cutter": {"kind": "cyl", "d": d2, "L": 20.0}, "r": r2, "h": h2 + 20.0, "zoff": h2}]}


def bull_cone(d1, r1, d2, a, l=10.0):
    """BullConeCutter's sub-cutters (compositecutter.cpp:349-366): the torus
    meets the cone tangentially where its slope is cot(a)."""
    h1 = r1 * math.sin(a)
    rad = math.sqrt(r1 * r1 - h1 * h1)
    rc = d1 / 2.0 - r1 + rad
    off = -(rc / math.tan(a) - (r1 - h1))
    height1 = r1 - h1
    h2 = (d2 / 2.0) / math.tan(a) + off
    return {"kind": "composite", "L": 30.0, "subs": [
        {"cutter": {"kind": "bull", "d": d1, "r": r1, "L": 1.0}, "r": rc, "h": height1, "zoff": 0.0},
        {"cutter": _cone(d2, a, l), "r": d2 / 2.0, "h": h2, "zoff": off},
        {"cutter": {"kind": "cyl", "d": d2, "L": 20.0}, "r": d2 / 2.0, "h": h2 + 20.0, "zoff": h2}]}


def cone_cone(d1, a1, d2, a2, l=10.0):
    """ConeConeCutter's sub-cutters (compositecutter.cpp:368-385): the outer
    cone continues the inner one at d1/2 (convex when a2 < a1)."""
    height1 = (d1 / 2.0) / math.tan(a1)
    off = -((d1 / 2.0) / math.tan(a2) - height1)
    h2 = (d2 / 2.0) / math.tan(a2) + off
    return {"kind": "composite", "L": 30.0, "subs": [
        {"cutter": _cone(d1, a1, l), "r": d1 / 2.0, "h": height1, "zoff": 0.0},
        {"cutter": _cone(d2, a2, l), "r": d2 / 2.0, "h": h2, "zoff": off},
        {"cutter": {"kind": "cyl", "d": d2, "L": 20.0}, "r": d2 / 2.0, "h": h2 + 20.0, "zoff": h2}]}

