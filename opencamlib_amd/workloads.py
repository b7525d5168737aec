"""Seeded synthetic inputs for drop-cutter and push-cutter.

All generators are deterministic (numpy PCG64 with an explicit seed) and return
float64 arrays, so every consumer (HIP path, CPU oracle, bench) reads identical
doubles.

Meshes are triangle soups: an (N, 9) float64 array, one row per triangle
(p0.xyz, p1.xyz, p2.xyz), the layout ``STLSurf::addTriangle`` receives.
"""
from __future__ import annotations

import math

import numpy as np


def heightfield(cells: int, extent: float, bumps: int, steps: int = 0, noise: float = 0.05,
                seed: int = 0, origin=(0.0, 0.0)) -> np.ndarray:
    """Heightfield mesh: `cells` x `cells` grid over [0, extent]^2, two triangles
    per cell split on the (i,j)->(i+1,j+1) diagonal.  z = sum of Gaussian bumps
    (amplitude U[5,20] mm, sigma U[8,25] mm, centres uniform) + axis-aligned steps
    (height U[2,10] mm added to vertices beyond a uniform position) + N(0, noise)."""
    rng = np.random.default_rng(seed)
    n = cells + 1
    g = np.linspace(0.0, extent, n)
    X, Y = np.meshgrid(g, g, indexing="xy")  # X[j, i] = g[i]
    Z = np.zeros_like(X)
    amp = rng.uniform(5.0, 20.0, bumps)
    sig = rng.uniform(8.0, 25.0, bumps)
    cx = rng.uniform(0.0, extent, bumps)
    cy = rng.uniform(0.0, extent, bumps)
    for a, s, x0, y0 in zip(amp, sig, cx, cy):
        Z += a * np.exp(-((X - x0) ** 2 + (Y - y0) ** 2) / (2.0 * s * s))
    for _ in range(steps):
        axis = rng.integers(0, 2)
        pos = rng.uniform(0.0, extent)
        hgt = rng.uniform(2.0, 10.0)
        if axis == 0:
            Z += np.where(X > pos, hgt, 0.0)
        else:
            Z += np.where(Y > pos, hgt, 0.0)
    Z += rng.normal(0.0, noise, Z.shape)
    X = X + origin[0]
    Y = Y + origin[1]
    P = np.stack([X, Y, Z], axis=-1)  # [j, i, 3]
    v00 = P[:-1, :-1].reshape(-1, 3)
    v10 = P[:-1, 1:].reshape(-1, 3)
    v11 = P[1:, 1:].reshape(-1, 3)
    v01 = P[1:, :-1].reshape(-1, 3)
    t1 = np.concatenate([v00, v10, v11], axis=1)
    t2 = np.concatenate([v00, v11, v01], axis=1)
    tris = np.stack([t1, t2], axis=1).reshape(-1, 9)
    return np.ascontiguousarray(tris)


def cl_grid(nx: int, ny: int, step: float, z0: float = -10.0, origin=(0.0, 0.0)) -> np.ndarray:
    """CL points (origin + step*i, origin + step*j, z0), row-major in j then i."""
    i = np.arange(nx, dtype=np.float64)
    j = np.arange(ny, dtype=np.float64)
    I, J = np.meshgrid(i, j, indexing="xy")
    x = origin[0] + step * I
    y = origin[1] + step * J
    return np.ascontiguousarray(np.stack([x.ravel(), y.ravel(), np.full(x.size, z0)], axis=1))


def cl_random(n: int, extent: float, z0: float = -10.0, seed: int = 0, origin=(0.0, 0.0)) -> np.ndarray:
    rng = np.random.default_rng(seed)
    xy = rng.uniform(0.0, extent, (n, 2))
    xy[:, 0] += origin[0]
    xy[:, 1] += origin[1]
    return np.ascontiguousarray(np.concatenate([xy, np.full((n, 1), z0)], axis=1))


def _icosahedron():
    t = (1.0 + math.sqrt(5.0)) / 2.0
    v = np.array([[-1, t, 0], [1, t, 0], [-1, -t, 0], [1, -t, 0],
                  [0, -1, t], [0, 1, t], [0, -1, -t], [0, 1, -t],
                  [t, 0, -1], [t, 0, 1], [-t, 0, -1], [-t, 0, 1]], dtype=np.float64)
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    f = np.array([[0, 11, 5], [0, 5, 1], [0, 1, 7], [0, 7, 10], [0, 10, 11],
                  [1, 5, 9], [5, 11, 4], [11, 10, 2], [10, 7, 6], [7, 1, 8],
                  [3, 9, 4], [3, 4, 2], [3, 2, 6], [3, 6, 8], [3, 8, 9],
                  [4, 9, 5], [2, 4, 11], [6, 2, 10], [8, 6, 7], [9, 8, 1]], dtype=np.int64)
    return v, f


def icosphere(level: int) -> np.ndarray:
    """Unit icosphere, 20*4^level triangles, (N, 9)."""
    v, f = _icosahedron()
    verts = [tuple(x) for x in v]
    index = {}

    def mid(a, b):
        key = (min(a, b), max(a, b))
        if key in index:
            return index[key]
        m = (np.asarray(verts[a]) + np.asarray(verts[b])) / 2.0
        m /= np.linalg.norm(m)
        verts.append(tuple(m))
        index[key] = len(verts) - 1
        return index[key]

    faces = [tuple(x) for x in f]
    for _ in range(level):
        nf = []
        for a, b, c in faces:
            ab, bc, ca = mid(a, b), mid(b, c), mid(c, a)
            nf += [(a, ab, ca), (b, bc, ab), (c, ca, bc), (ab, bc, ca)]
        faces = nf
    V = np.asarray(verts)
    F = np.asarray(faces)
    return np.ascontiguousarray(V[F].reshape(-1, 9))


def spheres_scene(count: int = 40, level: int = 4, extent: float = 200.0, seed: int = 2) -> np.ndarray:
    """`count` icospheres, radius U[5,20] mm, centres uniform in [0,extent]^2 at
    z = radius, plus a 2-triangle base plate at z = 0 over the domain."""
    rng = np.random.default_rng(seed)
    unit = icosphere(level).reshape(-1, 3, 3)
    parts = []
    for _ in range(count):
        r = rng.uniform(5.0, 20.0)
        c = np.array([rng.uniform(0.0, extent), rng.uniform(0.0, extent), r])
        parts.append((unit * r + c).reshape(-1, 9))
    plate = np.array([[0, 0, 0, extent, 0, 0, extent, extent, 0],
                      [0, 0, 0, extent, extent, 0, 0, extent, 0]], dtype=np.float64)
    parts.append(plate)
    return np.ascontiguousarray(np.concatenate(parts, axis=0))


def generate_range(start: float, end: float, n: int):
    """Waterline::generate_range (waterline.cpp:169-178): n+1 values, accumulated."""
    out = []
    d = (end - start) / float(n)
    v = start
    for _ in range(n + 1):
        out.append(v)
        v = v + d
    return out


def waterline_fibers(tris: np.ndarray, radius: float, z: float, nfib: int):
    """X and Y fibers at height z spanning the mesh bbox +- 2R
    (Waterline::init_fibers, waterline.cpp:143-165), nfib per direction.
    Returns (xfibers (nfib,6), yfibers (nfib,6))."""
    p = tris.reshape(-1, 3)
    minx, maxx = p[:, 0].min() - 2 * radius, p[:, 0].max() + 2 * radius
    miny, maxy = p[:, 1].min() - 2 * radius, p[:, 1].max() + 2 * radius
    ys = generate_range(miny, maxy, nfib - 1)
    xs = generate_range(minx, maxx, nfib - 1)
    xf = np.array([[minx, y, z, maxx, y, z] for y in ys], dtype=np.float64)
    yf = np.array([[x, miny, z, x, maxy, z] for x in xs], dtype=np.float64)
    return xf, yf


# ---------------------------------------------------------------- configs
def config0_mesh():
    return heightfield(64, 100.0, bumps=3, steps=0, seed=1000)


def config0_cl():
    return cl_grid(200, 200, 0.5, z0=-10.0)


def config1_mesh():
    return heightfield(512, 200.0, bumps=8, steps=20, seed=1001)


def config1_cl(n_side: int = 1000):
    return cl_grid(n_side, n_side, 0.2, z0=-10.0)


def config2_mesh():
    return spheres_scene(40, 4, 200.0, seed=1002)


def config2_levels(nlev: int = 20):
    return [1.0 + 29.0 * k / (nlev - 1) for k in range(nlev)]


def subsample(a: np.ndarray, n: int, seed: int) -> np.ndarray:
    """Seeded uniform subset of rows (sorted indices)."""
    if n >= a.shape[0]:
        return a
    rng = np.random.default_rng(seed)
    idx = np.sort(rng.choice(a.shape[0], size=n, replace=False))
    return np.ascontiguousarray(a[idx])


def config3_jobs(njobs: int = 64, seed: int = 1003, tri_range=(2_000, 500_000), cl_range=(1_000, 1_000_000)):
    """Uneven batched jobs: per job a heightfield as config 0 (3 bumps over
    100 x 100 mm) whose cell count is log-uniform so the mesh has
    tri_range triangles, a cutter uniform over Cyl/Ball/Bull/Cone/Composite
    (Cyl core d=6 + Ball rim d=10) with diameter U[3, 20] mm, and a CL grid
    of log-uniform size in cl_range at step U[0.2, 1] mm from a random
    origin.  Returns a list of dicts {tris, spec, cl}."""
    rng = np.random.default_rng(seed)
    jobs = []
    lo, hi = math.log(tri_range[0]), math.log(tri_range[1])
    clo, chi = math.log(cl_range[0]), math.log(cl_range[1])
    kinds = ["cyl", "ball", "bull", "cone", "composite"]
    for j in range(njobs):
        ntri = math.exp(rng.uniform(lo, hi))
        cells = max(2, int(round(math.sqrt(ntri / 2.0))))
        tris = heightfield(cells, 100.0, bumps=3, steps=0, seed=seed * 1000 + j)
        kind = kinds[int(rng.integers(0, 5))]
        d = float(rng.uniform(3.0, 20.0))
        if kind == "cyl":
            spec = {"kind": "cyl", "d": d, "L": 20.0}
        elif kind == "ball":
            spec = {"kind": "ball", "d": d, "L": 20.0}
        elif kind == "bull":
            spec = {"kind": "bull", "d": d, "r": 0.2 * d, "L": 20.0}
        elif kind == "cone":
            spec = {"kind": "cone", "d": d, "angle": math.pi / 4, "L": 10.0}
        else:
            spec = {"kind": "composite", "L": 20.0, "subs": [
                {"cutter": {"kind": "cyl", "d": 6.0, "L": 20.0}, "r": 3.0, "h": 0.0, "zoff": 0.0},
                {"cutter": {"kind": "ball", "d": 10.0, "L": 20.0}, "r": 5.0, "h": 4.0, "zoff": -1.0}]}
        n = int(math.exp(rng.uniform(clo, chi)))
        step = float(rng.uniform(0.2, 1.0))
        side = max(1, int(math.ceil(math.sqrt(n))))
        span = side * step
        ox = float(rng.uniform(0.0, max(0.0, 100.0 - span)))
        oy = float(rng.uniform(0.0, max(0.0, 100.0 - span)))
        cl = cl_grid(side, side, step, z0=-10.0, origin=(ox, oy))[:n]
        jobs.append({"tris": tris, "spec": spec, "cl": np.ascontiguousarray(cl)})
    return jobs


def random_polylines(n: int, extent: float, seed: int, nv=(3, 6), margin: float = 0.05):
    """n polylines of nv[0]..nv[1] vertices uniform in [margin, 1 - margin]
    x extent (the paths of the slope-driven refinement demo)."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        k = int(rng.integers(nv[0], nv[1] + 1))
        out.append(rng.uniform(margin * extent, (1.0 - margin) * extent, size=(k, 2)))
    return out
