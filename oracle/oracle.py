"""ctypes wrapper of the FP64 CPU restatement (oracle/ocl_oracle.c).

Test infrastructure only: imported by tests/, by ``__graft_entry__.smoke()``
and by the ``cpu_baseline`` leg of ``bench.py``, always as the checker or the
CPU baseline, never as the product path.
"""
from __future__ import annotations

import ctypes as C
import math
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "build", "libocl_oracle.so")

CYL, BALL, BULL, CONE, COMPOSITE = 0, 1, 2, 3, 4
MAX_SUB = 4
# parity diagnostics of the _diag entry points (ocl_oracle.h)
Q_COMP_EDGE, Q_CONE_TIP, Q_FACET_FIRST, Q_COMP_BELOW = 1, 2, 4, 8


class OCCutter(C.Structure):
    pass


OCCutter._fields_ = [
    ("kind", C.c_int),
    ("diameter", C.c_double), ("radius", C.c_double), ("length", C.c_double),
    ("xy_normal_length", C.c_double), ("normal_length", C.c_double), ("center_height", C.c_double),
    ("radius1", C.c_double), ("radius2", C.c_double),
    ("angle", C.c_double),
    ("nsub", C.c_int),
    ("sub", C.POINTER(OCCutter) * MAX_SUB),
    ("radiusvec", C.c_double * MAX_SUB), ("heightvec", C.c_double * MAX_SUB),
    ("zoffset", C.c_double * MAX_SUB),
]

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"oracle library missing: {LIB_PATH} (run __graft_entry__.build())")
        L = C.CDLL(LIB_PATH)
        dp = C.POINTER(C.c_double)
        ip = C.POINTER(C.c_int)
        lp = C.POINTER(C.c_long)
        cp = C.POINTER(OCCutter)
        for name, args in [("oc_cutter_cyl", [cp, C.c_double, C.c_double]),
                           ("oc_cutter_ball", [cp, C.c_double, C.c_double]),
                           ("oc_cutter_bull", [cp, C.c_double, C.c_double, C.c_double]),
                           ("oc_cutter_cone", [cp, C.c_double, C.c_double, C.c_double]),
                           ("oc_cutter_composite", [cp]),
                           ("oc_composite_add", [cp, cp, C.c_double, C.c_double, C.c_double])]:
            f = getattr(L, name)
            f.argtypes = args
            f.restype = None
        L.oc_mesh_create.argtypes = [dp, C.c_long]
        L.oc_mesh_create.restype = C.c_void_p
        L.oc_mesh_free.argtypes = [C.c_void_p]
        L.oc_mesh_free.restype = None
        L.oc_kd_rank.argtypes = [C.c_void_p, C.c_int, C.c_uint, ip]
        L.oc_kd_rank.restype = C.c_int
        L.oc_kd_search_drop.argtypes = [C.c_void_p, cp, C.c_double, C.c_double, C.c_double,
                                        C.c_uint, ip, C.c_long]
        L.oc_kd_search_drop.restype = C.c_long
        L.oc_kd_search_push.argtypes = [C.c_void_p, cp, C.c_int, dp, C.c_uint, ip, C.c_long]
        L.oc_kd_search_push.restype = C.c_long
        L.oc_batch_drop.argtypes = [C.c_void_p, cp, dp, C.c_long, C.c_int, C.c_int, C.c_uint,
                                    dp, dp, ip, ip, ip, lp]
        L.oc_batch_drop.restype = C.c_int
        L.oc_drop_triangle.argtypes = [cp, dp, C.c_double, C.c_double, dp, dp, ip, ip]
        L.oc_drop_triangle.restype = C.c_int
        L.oc_edge_drop.argtypes = [cp, dp, C.c_double, C.c_double, dp, dp]
        L.oc_edge_drop.restype = C.c_int
        L.oc_drop_triangle_diag.argtypes = [cp, dp, C.c_double, C.c_double, dp, dp, ip, ip, ip]
        L.oc_drop_triangle_diag.restype = C.c_int
        L.oc_batch_drop_diag.argtypes = [C.c_void_p, cp, dp, C.c_long, C.c_int, C.c_int, C.c_uint,
                                         dp, dp, ip, ip, ip, lp, ip]
        L.oc_batch_drop_diag.restype = C.c_int
        L.oc_batch_push.argtypes = [C.c_void_p, cp, C.c_int, dp, C.c_long, C.c_int, C.c_int, C.c_uint,
                                    C.c_long, lp, dp, dp, lp]
        L.oc_batch_push.restype = C.c_long
        L.oc_push_triangle.argtypes = [cp, dp, dp, dp]
        L.oc_push_triangle.restype = C.c_int
        L.oc_push_triangle_diag.argtypes = [cp, dp, dp, dp, ip]
        L.oc_push_triangle_diag.restype = C.c_int
        L.oc_fiber_union.argtypes = [dp, C.c_long, C.c_long, dp]
        L.oc_fiber_union.restype = C.c_long
        L.oc_drop_grad.argtypes = [cp, dp, C.c_int, C.c_int, C.c_double, C.c_double, dp, dp]
        L.oc_drop_grad.restype = C.c_int
        L.oc_batch_drop_grad.argtypes = [C.c_void_p, cp, dp, C.c_long, ip, ip, ip, dp, C.c_int, dp]
        L.oc_batch_drop_grad.restype = None
        _lib = L
    return _lib


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _ip(a):
    return a.ctypes.data_as(C.POINTER(C.c_int))


def _lp(a):
    return a.ctypes.data_as(C.POINTER(C.c_long))


class Cutter:
    """Oracle cutter; keeps sub-cutters alive for composites."""

    def __init__(self, kind, *p):
        self.c = OCCutter()
        self.subs = []
        L = lib()
        if kind == CYL:
            L.oc_cutter_cyl(C.byref(self.c), *p)
        elif kind == BALL:
            L.oc_cutter_ball(C.byref(self.c), *p)
        elif kind == BULL:
            L.oc_cutter_bull(C.byref(self.c), *p)
        elif kind == CONE:
            L.oc_cutter_cone(C.byref(self.c), *p)
        elif kind == COMPOSITE:
            L.oc_cutter_composite(C.byref(self.c))
        else:
            raise ValueError(kind)

    def add(self, sub: "Cutter", r, h, zoff):
        self.subs.append(sub)
        lib().oc_composite_add(C.byref(self.c), C.byref(sub.c), r, h, zoff)
        return self

    @property
    def radius(self):
        return self.c.radius

    @property
    def length(self):
        return self.c.length


def cutter_from_spec(spec: dict) -> Cutter:
    """Build an oracle cutter from the shared cutter-spec dictionaries used by
    tests and the bench (same keys as opencamlib_amd.cutter_spec)."""
    k = spec["kind"]
    if k == "cyl":
        return Cutter(CYL, spec["d"], spec["L"])
    if k == "ball":
        return Cutter(BALL, spec["d"], spec["L"])
    if k == "bull":
        return Cutter(BULL, spec["d"], spec["r"], spec["L"])
    if k == "cone":
        return Cutter(CONE, spec["d"], spec["angle"], spec["L"])
    if k == "composite":
        comp = Cutter(COMPOSITE)
        for s in spec["subs"]:
            comp.add(cutter_from_spec(s["cutter"]), s["r"], s["h"], s["zoff"])
        comp.c.length = spec["L"]
        return comp
    raise ValueError(k)


class Mesh:
    def __init__(self, tris):
        tris = np.ascontiguousarray(np.asarray(tris, dtype=np.float64).reshape(-1, 9))
        self.tris = tris
        self.h = lib().oc_mesh_create(_dp(tris), tris.shape[0])

    def __del__(self):
        if getattr(self, "h", None):
            lib().oc_mesh_free(self.h)
            self.h = None

    def __len__(self):
        return self.tris.shape[0]

    def kd_rank(self, dimset=0, bucket=1):
        r = np.empty(len(self), dtype=np.int32)
        lib().oc_kd_rank(self.h, dimset, bucket, _ip(r))
        return r

    def kd_search_drop(self, cutter: Cutter, x, y, z, bucket=1):
        cap = max(16, len(self))
        out = np.empty(cap, dtype=np.int32)
        n = lib().oc_kd_search_drop(self.h, C.byref(cutter.c), x, y, z, bucket, _ip(out), cap)
        return out[:n]

    def kd_search_push(self, cutter: Cutter, fib6, direction, bucket=1):
        cap = max(16, len(self))
        out = np.empty(cap, dtype=np.int32)
        f = np.ascontiguousarray(np.asarray(fib6, dtype=np.float64).reshape(6))
        n = lib().oc_kd_search_push(self.h, C.byref(cutter.c), direction, _dp(f), bucket, _ip(out), cap)
        return out[:n]

    def batch_drop(self, cutter: Cutter, cl_xyz, search_mode=1, nthreads=0, bucket=1, diag=False):
        """diag: also return `quirk`, the OC_Q_* bits per point (slower)."""
        cl = np.ascontiguousarray(np.asarray(cl_xyz, dtype=np.float64).reshape(-1, 3))
        n = cl.shape[0]
        z = np.empty(n)
        cc = np.empty((n, 3))
        ty = np.empty(n, dtype=np.int32)
        tri = np.empty(n, dtype=np.int32)
        feat = np.empty(n, dtype=np.int32)
        calls = np.zeros(1, dtype=np.int64)
        if diag:
            quirk = np.zeros(n, dtype=np.int32)
            lib().oc_batch_drop_diag(self.h, C.byref(cutter.c), _dp(cl), n, search_mode, nthreads, bucket,
                                     _dp(z), _dp(cc), _ip(ty), _ip(tri), _ip(feat), _lp(calls), _ip(quirk))
            return dict(z=z, cc=cc, type=ty, tri=tri, feat=feat, calls=int(calls[0]), quirk=quirk)
        lib().oc_batch_drop(self.h, C.byref(cutter.c), _dp(cl), n, search_mode, nthreads, bucket,
                            _dp(z), _dp(cc), _ip(ty), _ip(tri), _ip(feat), _lp(calls))
        return dict(z=z, cc=cc, type=ty, tri=tri, feat=feat, calls=int(calls[0]))

    def batch_drop_grad(self, cutter: Cutter, cl_xyz, res, nthreads=0):
        """Analytic FP64 gradient (n, 12) at every point of a batch_drop result
        res (tri, feat, type, cc): dz/dx, dz/dy, dz/dr, dz/d(p0, p1, p2)."""
        cl = np.ascontiguousarray(np.asarray(cl_xyz, dtype=np.float64).reshape(-1, 3))
        n = cl.shape[0]
        tri = np.ascontiguousarray(res["tri"], dtype=np.int32)
        feat = np.ascontiguousarray(res["feat"], dtype=np.int32)
        ty = np.ascontiguousarray(res["type"], dtype=np.int32)
        cc = np.ascontiguousarray(res["cc"], dtype=np.float64).reshape(-1, 3)
        out = np.empty((n, 12))
        lib().oc_batch_drop_grad(self.h, C.byref(cutter.c), _dp(cl), n, _ip(tri), _ip(feat), _ip(ty),
                                 _dp(cc), nthreads, _dp(out))
        return out

    def batch_push(self, cutter: Cutter, fibers, direction, search_mode=2, nthreads=0, bucket=1):
        fib = np.ascontiguousarray(np.asarray(fibers, dtype=np.float64).reshape(-1, 6))
        nf = fib.shape[0]
        cap = max(64, 16 * nf)
        calls = np.zeros(1, dtype=np.int64)
        while True:
            off = np.zeros(nf + 1, dtype=np.int64)
            lohi = np.empty(2 * cap)
            cc = np.empty(8 * cap)
            r = lib().oc_batch_push(self.h, C.byref(cutter.c), direction, _dp(fib), nf, search_mode,
                                    nthreads, bucket, cap, _lp(off), _dp(lohi), _dp(cc), _lp(calls))
            if r >= 0:
                break
            cap = -r
        return dict(offsets=off, lohi=lohi[:2 * r].reshape(-1, 2), cc=cc[:8 * r].reshape(-1, 8),
                    calls=int(calls[0]))


def drop_triangle(cutter: Cutter, tri9, x, y, z0=-1e9):
    """One triangle's MillingCutter::dropCutter; `quirk` holds the OC_Q_*
    bits of the parity diagnostics."""
    t = np.ascontiguousarray(np.asarray(tri9, dtype=np.float64).reshape(9))
    z = np.array([z0], dtype=np.float64)
    cc = np.zeros(3)
    ty = np.zeros(1, dtype=np.int32)
    ft = np.zeros(1, dtype=np.int32)
    q = np.zeros(1, dtype=np.int32)
    lifted = lib().oc_drop_triangle_diag(C.byref(cutter.c), _dp(t), x, y, _dp(z), _dp(cc), _ip(ty), _ip(ft),
                                         _ip(q))
    return dict(lifted=bool(lifted), z=float(z[0]), cc=cc, type=int(ty[0]), feat=int(ft[0]), quirk=int(q[0]))


def edge_drop(cutter: Cutter, p1p2, x, y, z0=-1e9):
    """One edge of MillingCutter::edgeDrop (simple cutters)."""
    e = np.ascontiguousarray(np.asarray(p1p2, dtype=np.float64).reshape(6))
    z = np.array([z0], dtype=np.float64)
    cc = np.zeros(3)
    lifted = lib().oc_edge_drop(C.byref(cutter.c), _dp(e), x, y, _dp(z), _dp(cc))
    return dict(lifted=bool(lifted), z=float(z[0]), cc=cc)


def push_triangle(cutter: Cutter, fib6, tri9, diag=False):
    """One (fiber, triangle) MillingCutter::pushCutter: (non-empty, out10),
    plus the OC_Q_* bits of the parity diagnostics with diag."""
    f = np.ascontiguousarray(np.asarray(fib6, dtype=np.float64).reshape(6))
    t = np.ascontiguousarray(np.asarray(tri9, dtype=np.float64).reshape(9))
    out = np.zeros(10)
    q = np.zeros(1, dtype=np.int32)
    ne = lib().oc_push_triangle_diag(C.byref(cutter.c), _dp(f), _dp(t), _dp(out), _ip(q))
    if diag:
        return bool(ne), out, int(q[0])
    return bool(ne), out


def drop_grad(cutter: Cutter, tri9, feat, cctype, x, y, cc3):
    """Analytic FP64 gradient of one drop result (see oc_drop_grad)."""
    t = np.ascontiguousarray(np.asarray(tri9, dtype=np.float64).reshape(9))
    c = np.ascontiguousarray(np.asarray(cc3, dtype=np.float64).reshape(3))
    out = np.zeros(12)
    lib().oc_drop_grad(C.byref(cutter.c), _dp(t), int(feat), int(cctype), float(x), float(y), _dp(c), _dp(out))
    return out


def fiber_union(ints10):
    a = np.ascontiguousarray(np.asarray(ints10, dtype=np.float64).reshape(-1, 10))
    out = np.empty((max(1, a.shape[0]), 10))
    n = lib().oc_fiber_union(_dp(a), a.shape[0], out.shape[0], _dp(out))
    return out[:n]


def cone_angle_deg(a):
    return a * 180.0 / math.pi
