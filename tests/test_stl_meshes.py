"""Real CAD meshes from the reference's stl/ folder (data files, copied into
tests/golden; sources and checksums in tests/golden/MANIFEST.txt).

CPU: the C ABI's STL reader (ocl_hip_stl_read, the reader behind
ocl::STLReader and ocl_hip_mesh_create_stl) equals an independent parse of
each file (binary records with numpy, text with Python float parsing rounded
to float32), and reads a binary file whose header starts with "solid" as
text, as the reference does (stlreader.cpp:63-66): ktoolcav.stl gives no
triangles.

GPU: each mesh, loaded straight from its file, through the batch
drop-cutter (every cutter kind, exact64 and mixed, against the FP64 oracle
to the bar of tests/test_drop_gpu.py) and the batch push-cutter (Bull and
Cone, X and Y fibers, against the oracle's KD-search candidates to the bar
of tests/test_push_gpu.py).
"""
import os
import struct

import numpy as np
import pytest

from cutters import scaled, ALL
from opencamlib_amd import capi, workloads as W
from oracle import oracle as O

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
# mesh: cutter diameter scaled to the part
MESHES = {"carpet1": 6.0, "mount_rush": 6.0, "sphere2": 3.0, "cone_on_side": 3.0, "wheel_in_box": 10.0,
          "spider": 2.0}


def independent_parse(path):
    data = open(path, "rb").read()
    if data[:5] == b"solid":
        tris, v = [], []
        for line in data.decode("latin-1").splitlines()[1:]:
            t = line.split()
            if not t:
                continue
            if t[0] == "vertex":
                xyz = [float(np.float32(float(x))) for x in t[1:4]]
                if len(v) < 3:
                    v.append(xyz)
                else:
                    v[2] = xyz
            elif t[0] == "facet":
                v = []
            elif t[0] == "endfacet" and len(v) == 3:
                tris.append(v[0] + v[1] + v[2])
        return np.array(tris, dtype=np.float64).reshape(-1, 9)
    (n,) = struct.unpack("<I", data[80:84])
    out = np.empty((n, 9))
    for i in range(n):
        out[i] = struct.unpack("<9f", data[84 + 50 * i + 12:84 + 50 * i + 48])
    return out


def path_of(name):
    return os.path.join(GOLDEN, name + ".stl")


@pytest.mark.parametrize("name", list(MESHES))
def test_stl_reader_equals_independent_parse(name):
    got = capi.stl_read(path_of(name))
    ref = independent_parse(path_of(name))
    assert got.shape == ref.shape and len(got) > 100
    assert np.array_equal(got, ref)


def test_binary_file_with_solid_header_reads_as_text():
    assert open(path_of("ktoolcav"), "rb").read(5) == b"solid"
    assert capi.stl_read(path_of("ktoolcav")).shape == (0, 9)
    with pytest.raises(capi.OclError, match="cannot read"):
        capi.stl_read(os.path.join(GOLDEN, "no_such_file.stl"))


def test_host_stl_reader_uses_the_same_triangles():
    from opencamlib_amd import ocl
    s = ocl.STLSurf()
    ocl.STLReader(path_of("mount_rush"), s)
    got = np.array([[c for p in (t.p[0], t.p[1], t.p[2]) for c in (p.x, p.y, p.z)] for t in s.getTriangles()])
    assert np.array_equal(got, capi.stl_read(path_of("mount_rush")))


def grid_over(tris, R, n=140):
    x0, x1 = tris[:, 0::3].min() - R, tris[:, 0::3].max() + R
    y0, y1 = tris[:, 1::3].min() - R, tris[:, 1::3].max() + R
    xs, ys = np.linspace(x0, x1, n), np.linspace(y0, y1, n)
    X, Y = np.meshgrid(xs, ys)
    z0 = tris[:, 2::3].min() - 10.0
    return np.stack([X.ravel(), Y.ravel(), np.full(X.size, z0)], axis=1)


@pytest.mark.gpu
@pytest.mark.parametrize("mode", ["exact64", "mixed"])
@pytest.mark.parametrize("kind", ["cyl", "ball", "bull", "cone", "cylball"])
@pytest.mark.parametrize("name", list(MESHES))
def test_stl_mesh_drop_matches_oracle(gpu, name, kind, mode):
    from test_drop_gpu import compare
    tris = capi.stl_read(path_of(name))
    spec = scaled(ALL[kind], MESHES[name])
    R = spec["d"] / 2.0 if "d" in spec else 5.0
    cl = grid_over(tris, R)
    dev = capi.DeviceMesh.from_stl(path_of(name))
    assert dev.ntri == len(tris)
    got = dev.drop(spec, cl, mode)
    dev.close()
    ref = O.Mesh(tris).batch_drop(O.cutter_from_spec(spec), cl, search_mode=1)
    compare(tris, spec, cl, got, ref, mode, label=name)


@pytest.mark.gpu
@pytest.mark.parametrize("precision", ["exact64", "mixed"])
@pytest.mark.parametrize("direction", ["x", "y"])
@pytest.mark.parametrize("kind", ["bull", "cone"])
@pytest.mark.parametrize("name", list(MESHES))
def test_stl_mesh_push_matches_oracle(gpu, name, kind, direction, precision):
    from test_push_gpu import compare_push
    tris = capi.stl_read(path_of(name))
    spec = scaled(ALL[kind], MESHES[name])
    z0, z1 = tris[:, 2::3].min(), tris[:, 2::3].max()
    fib = []
    for f in (0.2, 0.55, 0.85):
        xf, yf = W.waterline_fibers(tris, spec["d"] / 2.0, z0 + f * (z1 - z0), 100)
        fib.append(xf if direction == "x" else yf)
    fib = np.concatenate(fib)
    c, e, k = compare_push(tris, spec, fib, direction, name, precision=precision)
    assert c == 0 and e == 0 and k == 0
