"""The batched-jobs and refinement workloads at their full size (GPU).

Config 3: 64 uneven jobs (heightfields of 2k-500k triangles, 6.3 M in all;
Cyl/Ball/Bull/Cone/Composite cutters of diameter 3-20 mm; 1k-1M CL points
each, 10 M in all) in one persistent graph-captured batch, mixed precision.
Every job's outputs equal, bit for bit, its own ocl_hip_drop_device call,
and a seeded sample of 2,000 points per job matches the FP64 oracle to the
single-job bar (types exact, |dz| <= 1e-6 mm).

Config 4 demo: 500 random polylines (3-6 vertices, 164 m of path) on the
524,288-triangle heightfield, BallCutter d=10, refined by the slope-driven
AdaptivePathDropCutter to a 0.01 mm chord tolerance (first samples 1 mm
apart, minimum step 0.01 mm).  Every output interval's chord error is
measured by 15 dense device drops inside it.
"""
import numpy as np
import pytest

from cutters import ALL
from opencamlib_amd import capi, workloads as W
from oracle import oracle as O
from test_batch_gpu import FIELDS, DeviceJobs, same_bits

pytestmark = pytest.mark.gpu


def test_config3_full_batch_equals_single_calls_and_oracle(gpu):
    jobs = W.config3_jobs(64)
    ntri = sum(len(j["tris"]) for j in jobs)
    npts = sum(len(j["cl"]) for j in jobs)
    print(f"[config3] 64 jobs, {ntri} triangles, {npts} CL points", flush=True)
    d = DeviceJobs(jobs)
    b = capi.Batch(d.ctypes_jobs(), "mixed", {"use_graph": 1})
    b.run(want_stats=False)
    got = d.snapshot()
    b.close()
    ref = d.single("mixed")
    bad = [i for i in range(len(jobs)) if not same_bits([got[i]], [ref[i]])]
    print(f"[config3] jobs differing from their single call: {bad}", flush=True)
    assert not bad
    worst, ty_bad = 0.0, 0
    for i, j in enumerate(jobs):
        n = len(j["cl"])
        idx = np.sort(np.random.default_rng(100 + i).choice(n, min(2000, n), replace=False))
        r = O.Mesh(j["tris"]).batch_drop(O.cutter_from_spec(j["spec"]), j["cl"][idx], search_mode=1)
        worst = max(worst, float(np.max(np.abs(got[i]["z"][idx] - r["z"]))))
        ty_bad += int(np.sum(got[i]["type"][idx] != r["type"].astype(np.uint8)))
    print(f"[config3] oracle sample: max|dz|={worst:.3e} type mismatches={ty_bad}", flush=True)
    d.close()
    assert worst <= 1e-6 and ty_bad == 0


def refine_polylines(tris, spec, polylines, tol, sampling, min_step, z0=-20.0):
    """One AdaptivePathDropCutter over a path holding every polyline's
    segments; returns (points, per-interval chord error, interval length,
    levels, seconds)."""
    import time
    from opencamlib_amd import ocl
    s = ocl.STLSurf()
    for t in np.asarray(tris).reshape(-1, 9):
        s.addTriangle(ocl.Triangle(ocl.Point(*t[0:3]), ocl.Point(*t[3:6]), ocl.Point(*t[6:9])))
    cut = ocl.BallCutter(spec["d"], spec["L"])
    pth = ocl.Path()
    for poly in polylines:
        for a, b in zip(poly[:-1], poly[1:]):
            pth.append(ocl.Line(ocl.Point(a[0], a[1], z0), ocl.Point(b[0], b[1], z0)))
    apdc = ocl.AdaptivePathDropCutter()
    apdc.setSTL(s)
    apdc.setCutter(cut)
    apdc.setPath(pth)
    apdc.setSampling(sampling)
    apdc.setMinSampling(min_step)
    apdc.setChordTolerance(tol)
    t0 = time.perf_counter()
    apdc.run()
    secs = time.perf_counter() - t0
    p = np.array([[q.x, q.y, q.z] for q in apdc.getCLPoints()])
    # The output lists, span after span, the span's start and then the end of
    # each of its intervals, the last at the span's end.  Consecutive points
    # of one span form an interval; a span start follows the previous span's
    # end (the same point within a polyline, a jump between polylines).
    real = np.zeros(len(p) - 1, dtype=bool)
    i = 0
    for poly in polylines:
        for va, vb in zip(poly[:-1], poly[1:]):
            assert abs(p[i, 0] - va[0]) < 1e-9 and abs(p[i, 1] - va[1]) < 1e-9
            j = i + 1
            while abs(p[j, 0] - vb[0]) > 1e-9 or abs(p[j, 1] - vb[1]) > 1e-9:
                j += 1
            real[i:j] = True
            i = j + 1
    assert i == len(p)
    a, b = p[:-1][real], p[1:][real]
    tau = np.arange(1, 16) / 16.0
    probe = (a[:, None, :2] + (b - a)[:, None, :2] * tau[None, :, None]).reshape(-1, 2)
    cl = np.concatenate([probe, np.full((len(probe), 1), z0)], axis=1)
    dev = capi.DeviceMesh(tris)
    zt = dev.drop(spec, cl, "mixed", want_stats=False)["z"].reshape(len(a), 15)
    dev.close()
    chord = a[:, None, 2] + (b - a)[:, None, 2] * tau[None, :]
    err = np.max(np.abs(zt - chord), axis=1)
    return len(p), err, np.hypot(*(b - a)[:, :2].T), apdc.getLevels(), secs


def test_config4_refinement_demo_500_polylines(gpu):
    tol, min_step = 0.01, 0.01
    tris = W.config1_mesh()
    polys = W.random_polylines(500, 200.0, seed=44)
    total_len = sum(np.sum(np.hypot(*np.diff(p, axis=0).T)) for p in polys)
    npts, err, length, levels, secs = refine_polylines(tris, ALL["ball"], polys, tol, 1.0, min_step)
    long_ = length > 2 * min_step
    frac = float(np.mean(err[long_] <= tol))
    print(f"[config4] 500 polylines, {total_len:.0f} mm of path: points={npts} intervals={len(err)} "
          f"levels={levels} refine {secs:.2f} s; chord error max={err.max():.4f} p99={np.percentile(err, 99):.4f} "
          f"within tol (intervals > 2 min steps)={frac:.5f}; uniform at the minimum step would need "
          f"{int(total_len / min_step)} points", flush=True)
    assert frac >= 0.995 and err[long_].max() <= 1.5 * tol
    assert npts < total_len / min_step / 5
