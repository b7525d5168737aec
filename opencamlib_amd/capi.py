"""ctypes binding of the C ABI in include/ocl_hip.h (libocl_hip.so).

This is the thin layer the Python package, the tests and the bench use to
reach the HIP kernels.  It fails loudly when the native library is missing:
there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from ._runtime import preload_hip_runtime

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "libocl_hip.so")

CUTTER_TYPES = {"cyl": 0, "ball": 1, "bull": 2, "cone": 3, "composite": 4}
PREC = {"exact64": 0, "fp64": 1, "fp32": 2, "mixed": 3}
PUSH_EXACT_FILTER = 0x100

# CCType names (src/geo/ccpoint.hpp:33-39)
CCTYPE_NAMES = ["NONE", "VERTEX", "VERTEX_CYL", "EDGE", "EDGE_HORIZ", "EDGE_SHAFT", "EDGE_HORIZ_CYL",
                "EDGE_HORIZ_TOR", "EDGE_BALL", "EDGE_POS", "EDGE_NEG", "EDGE_CYL", "EDGE_CONE",
                "EDGE_CONE_BASE", "FACET", "FACET_TIP", "FACET_CYL", "ERROR"]


class CutterSimple(C.Structure):
    _fields_ = [("type", C.c_int32), ("diameter", C.c_double), ("length", C.c_double),
                ("corner_radius", C.c_double), ("angle", C.c_double)]


class CutterDesc(C.Structure):
    _fields_ = [("c", CutterSimple), ("nsub", C.c_int32), ("sub", CutterSimple * 4),
                ("radius_limit", C.c_double * 4), ("height_limit", C.c_double * 4),
                ("zoffset", C.c_double * 4)]


class Stats(C.Structure):
    _fields_ = [("points", C.c_uint64), ("pairs", C.c_uint64), ("escalated", C.c_uint64),
                ("escalated_exact", C.c_uint64),
                ("brent_iters", C.c_uint64), ("intervals", C.c_uint64), ("kernel_ms", C.c_double),
                ("select_ms", C.c_double), ("wave_evals", C.c_uint64), ("skipped", C.c_uint64)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class DropJob(C.Structure):
    _fields_ = [("mesh", C.c_void_p), ("cutter", CutterDesc), ("d_cl_xyz", C.c_void_p), ("n", C.c_size_t),
                ("d_z_out", C.c_void_p), ("d_cc_xyz_out", C.c_void_p), ("d_cc_type_out", C.c_void_p),
                ("d_tri_out", C.c_void_p), ("d_feat_out", C.c_void_p)]


class Tune(C.Structure):
    _fields_ = [("hilbert_bits", C.c_int32), ("blocks_per_cu", C.c_int32), ("use_graph", C.c_int32),
                ("per_job_launch", C.c_int32)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"native library missing: {LIB_PATH}; run __graft_entry__.build()")
        preload_hip_runtime()
        L = C.CDLL(LIB_PATH)
        vp, dp, sz = C.c_void_p, C.POINTER(C.c_double), C.c_size_t
        L.ocl_hip_mesh_create.argtypes = [dp, sz, C.c_uint, C.POINTER(C.c_void_p)]
        L.ocl_hip_mesh_create.restype = C.c_int
        L.ocl_hip_mesh_destroy.argtypes = [vp]
        L.ocl_hip_mesh_destroy.restype = C.c_int
        L.ocl_hip_mesh_size.argtypes = [vp]
        L.ocl_hip_mesh_size.restype = sz
        L.ocl_hip_drop_batch.argtypes = [vp, C.POINTER(CutterDesc), dp, sz, C.c_int, dp, dp,
                                         C.POINTER(C.c_uint8), C.POINTER(C.c_uint32), C.POINTER(C.c_uint8),
                                         C.POINTER(Stats), vp]
        L.ocl_hip_drop_batch.restype = C.c_int
        L.ocl_hip_drop_device.argtypes = [vp, C.POINTER(CutterDesc), vp, sz, C.c_int, vp, vp, vp, vp, vp,
                                          C.POINTER(Stats), vp]
        L.ocl_hip_drop_device.restype = C.c_int
        u64p = C.POINTER(C.c_uint64)
        L.ocl_hip_push_batch.argtypes = [vp, C.POINTER(CutterDesc), C.c_int, dp, sz, C.c_int, u64p, u64p,
                                         C.POINTER(Stats), vp]
        L.ocl_hip_push_batch.restype = C.c_int
        L.ocl_hip_push_fetch.argtypes = [vp, dp, dp, vp]
        L.ocl_hip_push_fetch.restype = C.c_int
        L.ocl_hip_push_xy.argtypes = [vp, C.POINTER(CutterDesc), dp, sz, dp, sz, C.c_int, u64p, u64p, u64p,
                                      C.POINTER(Stats), vp]
        L.ocl_hip_push_xy.restype = C.c_int
        L.ocl_hip_push_fetch_dir.argtypes = [vp, C.c_int, dp, dp, vp]
        L.ocl_hip_push_fetch_dir.restype = C.c_int
        L.ocl_hip_stl_read.argtypes = [C.c_char_p, dp, sz, C.POINTER(sz)]
        L.ocl_hip_stl_read.restype = C.c_int
        L.ocl_hip_mesh_create_stl.argtypes = [C.c_char_p, C.c_uint, C.POINTER(vp)]
        L.ocl_hip_mesh_create_stl.restype = C.c_int
        L.ocl_hip_drop_device_noalloc.argtypes = [vp, C.POINTER(CutterDesc), vp, sz, C.c_int, vp, vp, vp, vp, vp,
                                                  vp]
        L.ocl_hip_drop_device_noalloc.restype = C.c_int
        L.ocl_hip_drop_backward.argtypes = [vp, C.POINTER(CutterDesc), vp, sz, vp, vp, vp, vp, vp, C.c_int, vp, vp,
                                            vp, vp]
        L.ocl_hip_drop_backward.restype = C.c_int
        L.ocl_hip_drop_slopes.argtypes = [vp, C.POINTER(CutterDesc), dp, sz, C.c_int, dp, dp, dp,
                                          C.POINTER(C.c_uint8), C.POINTER(Stats), vp]
        L.ocl_hip_drop_slopes.restype = C.c_int
        L.ocl_hip_batch_create.argtypes = [C.POINTER(DropJob), sz, C.c_int, C.POINTER(Tune), C.POINTER(C.c_void_p)]
        L.ocl_hip_batch_create.restype = C.c_int
        L.ocl_hip_batch_run.argtypes = [vp, C.POINTER(Stats), vp]
        L.ocl_hip_batch_run.restype = C.c_int
        L.ocl_hip_batch_job_stats.argtypes = [vp, C.c_size_t, C.POINTER(Stats)]
        L.ocl_hip_batch_job_stats.restype = C.c_int
        L.ocl_hip_batch_destroy.argtypes = [vp]
        L.ocl_hip_batch_destroy.restype = C.c_int
        L.ocl_hip_batch_autotune.argtypes = [C.POINTER(DropJob), sz, C.c_int, C.c_int, C.POINTER(Tune),
                                             C.POINTER(C.c_double)]
        L.ocl_hip_batch_autotune.restype = C.c_int
        L.ocl_hip_kd_order.argtypes = [dp, sz, C.c_int, C.c_uint, C.POINTER(C.c_uint32), dp]
        L.ocl_hip_kd_order.restype = C.c_int
        L.ocl_hip_last_error.argtypes = []
        L.ocl_hip_last_error.restype = C.c_char_p
        _lib = L
    return _lib


class OclError(RuntimeError):
    pass


def check(rc):
    if rc != 0:
        raise OclError(f"ocl_hip error {rc}: {lib().ocl_hip_last_error().decode()}")


def _simple(spec) -> CutterSimple:
    s = CutterSimple()
    s.type = CUTTER_TYPES[spec["kind"]]
    s.diameter = float(spec.get("d", 0.0))
    s.length = float(spec.get("L", 0.0))
    s.corner_radius = float(spec.get("r", 0.0))
    s.angle = float(spec.get("angle", 0.0))
    return s


def cutter_desc(spec: dict) -> CutterDesc:
    """Descriptor from a cutter spec dict: {"kind": "cyl"|"ball"|"bull"|"cone", "d",
    "L", "r" (bull corner radius), "angle" (cone half-angle, rad)} or
    {"kind": "composite", "L", "subs": [{"cutter": spec, "r", "h", "zoff"}, ...]}."""
    d = CutterDesc()
    if spec["kind"] == "composite":
        d.c.type = CUTTER_TYPES["composite"]
        d.c.length = float(spec["L"])
        d.nsub = len(spec["subs"])
        for i, s in enumerate(spec["subs"]):
            d.sub[i] = _simple(s["cutter"])
            d.radius_limit[i] = float(s["r"])
            d.height_limit[i] = float(s["h"])
            d.zoffset[i] = float(s["zoff"])
    else:
        d.c = _simple(spec)
    return d


def _ptr(a, t):
    return a.ctypes.data_as(C.POINTER(t))


DIMSETS = {"xy": 0, "yz": 1, "xz": 2}


def kd_order(tris, dimset="xy", bucket: int = 1, with_accept: bool = False):
    """order[rank] = triangle index in the reference KD-tree's leaf order
    (host computation, no device needed).  with_accept: also the (ntri, 4)
    acceptance bounds per rank (A0, B0, A1, B1) of the KD search."""
    t = np.ascontiguousarray(np.asarray(tris, dtype=np.float64).reshape(-1, 9))
    out = np.empty(t.shape[0], dtype=np.uint32)
    acc = np.empty((t.shape[0], 4)) if with_accept else None
    check(lib().ocl_hip_kd_order(_ptr(t, C.c_double), t.shape[0], DIMSETS[dimset], bucket,
                                 _ptr(out, C.c_uint32), _ptr(acc, C.c_double) if with_accept else None))
    return (out, acc) if with_accept else out


def stl_read(path) -> np.ndarray:
    """Triangles (n, 9) of a binary or text STL file, float32 coordinates
    promoted to double (ocl_hip_stl_read; host only)."""
    p = os.fsencode(path)
    n = C.c_size_t(0)
    check(lib().ocl_hip_stl_read(p, None, 0, C.byref(n)))
    out = np.empty((n.value, 9))
    if n.value:
        check(lib().ocl_hip_stl_read(p, _ptr(out, C.c_double), n.value, C.byref(n)))
    return out


class DeviceMesh:
    """Owns an ocl_mesh handle (device copy + box hierarchy in KD order)."""

    def __init__(self, tris, bucket: int = 1):
        t = np.ascontiguousarray(np.asarray(tris, dtype=np.float64).reshape(-1, 9))
        h = C.c_void_p()
        check(lib().ocl_hip_mesh_create(_ptr(t, C.c_double), t.shape[0], bucket, C.byref(h)))
        self.h = h
        self.ntri = t.shape[0]

    @classmethod
    def from_stl(cls, path, bucket: int = 1):
        """A device mesh straight from an STL file (ocl_hip_mesh_create_stl)."""
        self = cls.__new__(cls)
        h = C.c_void_p()
        check(lib().ocl_hip_mesh_create_stl(os.fsencode(path), bucket, C.byref(h)))
        self.h = h
        self.ntri = int(lib().ocl_hip_mesh_size(h))
        return self

    def close(self):
        """Frees the handle; raises OclError (and keeps it) while a batch
        still uses the mesh."""
        if getattr(self, "h", None):
            check(lib().ocl_hip_mesh_destroy(self.h))
            self.h = None

    def __del__(self):
        # at interpreter exit the module's globals may already be gone
        try:
            self.close()
        except Exception:
            pass

    def drop(self, spec: dict, cl_xyz, precision="exact64", want_stats=True):
        """Batch drop of host CL points (n, 3).  Returns z, cc (n, 3), type
        (CCType), tri (input index, -1 where nothing was touched), feat."""
        cl = np.ascontiguousarray(np.asarray(cl_xyz, dtype=np.float64).reshape(-1, 3))
        n = cl.shape[0]
        z = np.empty(n)
        cc = np.empty((n, 3))
        ty = np.empty(n, dtype=np.uint8)
        tri = np.empty(n, dtype=np.uint32)
        ft = np.empty(n, dtype=np.uint8)
        st = Stats()
        desc = cutter_desc(spec)
        check(lib().ocl_hip_drop_batch(self.h, C.byref(desc), _ptr(cl, C.c_double), n, PREC[precision],
                                       _ptr(z, C.c_double), _ptr(cc, C.c_double), _ptr(ty, C.c_uint8),
                                       _ptr(tri, C.c_uint32), _ptr(ft, C.c_uint8),
                                       C.byref(st) if want_stats else None, None))
        tri64 = tri.astype(np.int64)
        tri64[tri == 0xFFFFFFFF] = -1
        return dict(z=z, cc=cc, type=ty, tri=tri64, feat=ft, stats=st.as_dict())

    def push(self, spec: dict, fibers, direction, precision="exact64", want_stats=True, exact_filter=False):
        """Batch push-cutter of host fibers (nf, 6) along X (direction 0 / "x")
        or Y (1 / "y").  Returns offsets (nf+1), lohi (rows, 2), cc (rows, 8)
        and stats; fiber i owns rows offsets[i]:offsets[i+1], sorted by lower.
        exact_filter: push only triangles whose boxes overlap the search box
        (the reference pushes its whole KD-search superset)."""
        d = {"x": 0, "y": 1}.get(direction, direction)
        fib = np.ascontiguousarray(np.asarray(fibers, dtype=np.float64).reshape(-1, 6))
        nf = fib.shape[0]
        offs = np.zeros(nf + 1, dtype=np.uint64)
        total = C.c_uint64(0)
        st = Stats()
        desc = cutter_desc(spec)
        prec = PREC[precision] | (PUSH_EXACT_FILTER if exact_filter else 0)
        check(lib().ocl_hip_push_batch(self.h, C.byref(desc), int(d), _ptr(fib, C.c_double), nf, prec,
                                       _ptr(offs, C.c_uint64), C.byref(total),
                                       C.byref(st) if want_stats else None, None))
        n = int(total.value)
        lohi = np.empty((n, 2))
        cc = np.empty((n, 8))
        if n:
            check(lib().ocl_hip_push_fetch(self.h, _ptr(lohi, C.c_double), _ptr(cc, C.c_double), None))
        return dict(offsets=offs.astype(np.int64), lohi=lohi, cc=cc, stats=st.as_dict())

    def push_xy(self, spec: dict, xfibers, yfibers, precision="exact64", want_stats=True, exact_filter=False):
        """X and Y fibers in one launch (ocl_hip_push_xy).  Returns {"x": ...,
        "y": ..., "stats": ...}, each direction as push() returns it."""
        fx = np.ascontiguousarray(np.asarray(xfibers, dtype=np.float64).reshape(-1, 6))
        fy = np.ascontiguousarray(np.asarray(yfibers, dtype=np.float64).reshape(-1, 6))
        ox = np.zeros(fx.shape[0] + 1, dtype=np.uint64)
        oy = np.zeros(fy.shape[0] + 1, dtype=np.uint64)
        tot = np.zeros(2, dtype=np.uint64)
        st = Stats()
        desc = cutter_desc(spec)
        prec = PREC[precision] | (PUSH_EXACT_FILTER if exact_filter else 0)
        check(lib().ocl_hip_push_xy(self.h, C.byref(desc), _ptr(fx, C.c_double), fx.shape[0], _ptr(fy, C.c_double),
                                    fy.shape[0], prec, _ptr(ox, C.c_uint64), _ptr(oy, C.c_uint64),
                                    _ptr(tot, C.c_uint64), C.byref(st) if want_stats else None, None))
        out = {"stats": st.as_dict()}
        for d, offs in ((0, ox), (1, oy)):
            n = int(tot[d])
            lohi = np.empty((n, 2))
            cc = np.empty((n, 8))
            if n:
                check(lib().ocl_hip_push_fetch_dir(self.h, d, _ptr(lohi, C.c_double), _ptr(cc, C.c_double), None))
            out["xy"[d]] = dict(offsets=offs.astype(np.int64), lohi=lohi, cc=cc)
        return out

    def drop_device(self, spec: dict, d_cl: int, n: int, precision, d_z: int, d_cc=0, d_type=0, d_tri=0,
                    d_feat=0, stream=None, stats=False):
        """Asynchronous drop on device pointers (ints, e.g. torch data_ptr())."""
        st = Stats()
        desc = cutter_desc(spec)
        check(lib().ocl_hip_drop_device(self.h, C.byref(desc), d_cl, n, PREC[precision], d_z,
                                        d_cc or None, d_type or None, d_tri or None, d_feat or None,
                                        C.byref(st) if stats else None, stream))
        return st.as_dict() if stats else None

    def drop_device_noalloc(self, spec: dict, d_cl: int, n: int, precision, d_z: int, d_cc=0, d_type=0, d_tri=0,
                            d_feat=0, stream=None):
        """ocl_hip_drop_device_noalloc: drop_device without device allocation
        (at most 524288 points; the mesh must already hold this cutter's
        selection records)."""
        desc = cutter_desc(spec)
        check(lib().ocl_hip_drop_device_noalloc(self.h, C.byref(desc), d_cl, n, PREC[precision], d_z, d_cc or None,
                                                d_type or None, d_tri or None, d_feat or None, stream))

    def drop_backward(self, spec: dict, d_cl: int, n: int, d_cc: int, d_type: int, d_tri: int, d_feat: int,
                      d_grad_z=0, precision="mixed", d_dz=0, d_grad_tri=0, d_grad_r=0, stream=None):
        """Asynchronous backward on device pointers (see ocl_hip_drop_backward):
        from a forward call's cc/type/tri/feat arrays, writes dz/dx, dz/dy,
        dz/dr per point to d_dz (n x 3) and adds the upstream-weighted vertex
        and radius sums to d_grad_tri (ntri x 9) and d_grad_r."""
        desc = cutter_desc(spec)
        check(lib().ocl_hip_drop_backward(self.h, C.byref(desc), d_cl, n, d_cc, d_type, d_tri, d_feat,
                                          d_grad_z or None, PREC[precision], d_dz or None, d_grad_tri or None,
                                          d_grad_r or None, stream))


def make_jobs(items):
    """ctypes array of DropJob from dicts {mesh: DeviceMesh, spec, cl, n, z, cc,
    type, tri, feat} whose arrays are device pointers (ints).  The array keeps
    its DeviceMesh objects alive (`meshes`), and so does every Batch built
    from it."""
    arr = (DropJob * len(items))()
    arr.meshes = [it["mesh"] for it in items]
    for i, it in enumerate(items):
        j = arr[i]
        j.mesh = it["mesh"].h
        j.cutter = cutter_desc(it["spec"])
        j.d_cl_xyz = it["cl"]
        j.n = it["n"]
        j.d_z_out = it["z"]
        j.d_cc_xyz_out = it.get("cc") or None
        j.d_cc_type_out = it.get("type") or None
        j.d_tri_out = it.get("tri") or None
        j.d_feat_out = it.get("feat") or None
    return arr


class Batch:
    """ocl_batch handle (see ocl_hip_batch_create)."""

    def __init__(self, jobs, precision="mixed", tune=None):
        t = Tune(**(tune or {}))
        h = C.c_void_p()
        check(lib().ocl_hip_batch_create(jobs, len(jobs), PREC[precision], C.byref(t), C.byref(h)))
        self.h = h
        self.jobs = jobs
        # a mesh must outlive every batch that uses it (ocl_hip_mesh_destroy
        # refuses while one is alive)
        self.meshes = list(getattr(jobs, "meshes", []))

    def run(self, stream=None, want_stats=True):
        st = Stats()
        check(lib().ocl_hip_batch_run(self.h, C.byref(st) if want_stats else None, stream))
        return st.as_dict() if want_stats else None

    def job_stats(self, j):
        """Counters of job j in the last run (points, pairs, escalated,
        wave_evals)."""
        st = Stats()
        check(lib().ocl_hip_batch_job_stats(self.h, j, C.byref(st)))
        return st.as_dict()

    def close(self):
        if getattr(self, "h", None):
            lib().ocl_hip_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def autotune(jobs, precision="mixed", trials=3):
    best = Tune()
    ms = C.c_double(0)
    check(lib().ocl_hip_batch_autotune(jobs, len(jobs), PREC[precision], trials, C.byref(best), C.byref(ms)))
    return best.as_dict(), ms.value
