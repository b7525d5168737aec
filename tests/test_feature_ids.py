"""Feature numbering of a triangle soup (CPU): the vertex and edge IDs that
let the FP32 selection tell a second copy of the winning contact (the same
vertex or edge seen from a neighbouring triangle) from a different contact.

Pinned by what the mesh fixes: two corners share a vertex ID exactly when
their coordinates are equal (checked against numpy's row-unique), a closed
surface uses every edge twice and in opposite directions, and its counts
satisfy Euler's V - E + F = 2 per connected genus-0 shell (sphere2.stl holds
two spheres)."""
import ctypes as C
import os

import numpy as np
import pytest

from opencamlib_amd import capi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")
REV = 0x40000000


@pytest.fixture(scope="module")
def ids_of():
    L = C.CDLL(os.path.join(ROOT, "opencamlib_amd", "lib", "libocl_selcheck.so"))
    L.ocl_selcheck_feature_ids.argtypes = [C.c_void_p, C.c_size_t, C.c_void_p]

    def run(tris):
        t = np.ascontiguousarray(np.asarray(tris, dtype=np.float64).reshape(-1, 9))
        out = np.empty(6 * len(t), dtype=np.uint32)
        assert L.ocl_selcheck_feature_ids(t.ctypes.data, len(t), out.ctypes.data) == 0
        return out.reshape(-1, 6).astype(np.int64)
    return run


def split(ids):
    return ids[:, :3], ids[:, 3:] & ~REV, (ids[:, 3:] & REV) != 0


def test_two_triangles_sharing_an_edge(ids_of):
    a, b, c, d = (0, 0, 1), (1, 0, 1), (0, 1, 1), (1, 1, 2)
    ids = ids_of([a + b + c, b + d + c])
    v, e, rev = split(ids)
    assert v[0, 1] == v[1, 0] and v[0, 2] == v[1, 2]          # b and c shared
    assert len(np.unique(v)) == 4 and e.min() >= 4             # edges after vertices
    # edge (b, c) is edge 1 of the first triangle and edge 2 of the second,
    # run in opposite directions; the other four edges are distinct
    assert e[0, 1] == e[1, 2] and rev[0, 1] != rev[1, 2]
    assert len(np.unique(e)) == 5


def test_nan_coordinates_share_nothing(ids_of):
    t = np.array([[0, 0, 0, 1, 0, 0, 0, 1, 0], [1, 0, 0, np.nan, 1, 0, 0, 1, 0]], dtype=float)
    v, e, _ = split(ids_of(t))
    assert len(np.unique(v)) == 6 and len(np.unique(e)) == 6


@pytest.mark.parametrize("name", ["mount_rush", "spider"])
def test_vertex_ids_equal_exactly_when_coordinates_do(ids_of, name):
    t = capi.stl_read(os.path.join(GOLDEN, name + ".stl"))
    v, _, _ = split(ids_of(t))
    corners = t.reshape(-1, 3)
    _, inv = np.unique(corners, axis=0, return_inverse=True)
    # same partition of the 3n corners
    pairs = np.unique(np.stack([v.ravel(), inv.ravel()], axis=1), axis=0)
    assert len(pairs) == len(np.unique(inv)) == len(np.unique(v))


def test_closed_spheres_euler_and_edge_directions(ids_of):
    t = capi.stl_read(os.path.join(GOLDEN, "sphere2.stl"))
    v, e, rev = split(ids_of(t))
    ev, er = e.ravel(), rev.ravel()
    order = np.argsort(ev, kind="stable")
    counts = np.bincount(ev - ev.min())
    assert set(np.unique(counts[counts > 0])) == {2}            # every edge twice
    assert np.all(er[order][0::2] != er[order][1::2])           # once each way
    V, E, F = len(np.unique(v)), len(np.unique(ev)), len(t)
    assert V - E + F == 4                                       # two spheres
