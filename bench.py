#!/usr/bin/env python3
"""Batch drop-cutter throughput on one MI355X.

Workload: the 524,288-triangle heightfield (512x512 cells over 200x200 mm,
8 Gaussian bumps + 20 axis-aligned steps, seed 1001) with a 1000x1000 CL grid
at 0.2 mm (z0 = -10), BallCutter d=10 mm L=20 mm (``--cutter bull`` for the
BullCutter d=10 r=2 run).  One step = one batch drop of all 1,000,000 CL
points with the CL array already in HBM, through the C ABI entry point
``ocl_hip_drop_device`` in mixed precision (FP32 selection, FP64 escalation
and FP64 refinement of the winning contact).

Beside it, in the same run: the FP64-only and FP32-only device paths, the
reference-order exact FP64 device path, the accuracy of each against the FP64
oracle on a seeded sample, and the oracle itself timed on the host cores on a
bounded sample (the ``cpu_baseline``).  Unless ``--no-extra``: the same
mesh and grid with the BullCutter (the configuration's second run), the backward
pass on 1M random points of the same mesh, the push-cutter on the 80,000
waterline fibers of the sphere scene (Bull and Cone, X and Y fibers in one
launch), 64 uneven batched jobs (persistent graph launch vs per-job
launches) and the slope-driven refinement of 500 polylines on the same mesh.
Every section with a hot kernel carries a roofline object.  Prints one JSON
line.
"""
from __future__ import annotations

import argparse
import json
import os
import subprocess
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = ("drop-cutter CL pts/s @524k tris: FP32+FP64-fix vs FP64 GPU, |Δz|≤1µm; %VALU peak")
# MI355X published vector peaks (TFLOP/s): FP32 with packed math, FP32 without
# (packed FP32 instructions are not used by these kernels), FP64 vector.
PEAK_FP32 = 157.3
PEAK_FP32_UNPACKED = 78.6
PEAK_FP64 = 78.6
# Algorithmic FP32 work per evaluated (CL point, triangle) pair of the
# selection's contact function (facet + 3 vertices + 3 edges with their error
# bounds; add/sub/mul = 1, fma = 2, sqrt/rsqrt/rcp/div = 1), counted by
# running that code on the host with a counting float type
# (tests/flops/flop_drop.cpp; tests/test_flop_counts.py keeps these equal).
F_PAIR = {"cyl": 102, "ball": 110, "bull": 480, "cone": 165}
# FP64 work per (fiber, candidate triangle) pair of the push contact code,
# the same rule, averaged over the KD-superset pairs of 40 seeded X fibers of
# the push workload (tests/flops/flop_push.cpp); BullCutter includes its
# edges' Brent solves (15.2 iterations per pair on that sample).
F_PUSH = {"bull": 2430, "cone": 1244}
# The same count over the pairs the mixed-precision push keeps (the sample's
# pairs less those the containment bound leaves out, replayed in rank order
# in batches of 64 as the kernel takes them).
F_PUSH_KEPT = {"bull": 3337, "cone": 1344}

SPECS = {
    "ball": {"kind": "ball", "d": 10.0, "L": 20.0},
    "bull": {"kind": "bull", "d": 10.0, "r": 2.0, "L": 20.0},
    "cyl": {"kind": "cyl", "d": 6.0, "L": 20.0},
    "cone": {"kind": "cone", "d": 10.0, "angle": np.pi / 4, "L": 10.0},
}


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def host_cpu():
    """The host CPU as lscpu reports it, and the CPUs this process may run on."""
    info = {}
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=20).stdout
        keys = {"Model name": "model", "Socket(s)": "sockets", "Core(s) per socket": "cores_per_socket",
                "Thread(s) per core": "threads_per_core", "CPU(s)": "logical_cpus"}
        for line in out.splitlines():
            k, _, v = line.partition(":")
            if k.strip() in keys:
                info[keys[k.strip()]] = v.strip()
        if "sockets" in info and "cores_per_socket" in info:
            info["physical_cores"] = int(info["sockets"]) * int(info["cores_per_socket"])
    except (OSError, ValueError, subprocess.SubprocessError):
        pass
    info["cpus_available_to_this_process"] = len(os.sched_getaffinity(0))
    return info


# The reference itself, timed by the survey in the build container (it is not
# built here: its build needs Boost); beside the port's rate for scale.
REFERENCE_MEASURED = {
    "value": 136.0, "unit": "CL pts/s", "cores": 8, "host": "8-core Intel Xeon @ 2.70 GHz, 1 thread per core",
    "workload": "2,500 CL points spread over a 524,288-triangle heightfield (8 bumps, no steps), BallCutter d=10",
    "source": "survey measurement of the reference's own build (not on the GPU box)"}


def time_mode(mesh, spec, d_cl, n, mode, outs, steps, warmup, torch):
    """Runs warmup steps, then `steps` timed steps back to back (no host
    synchronisation inside), then `steps` more with per-call statistics, whose
    HIP events, recorded on the stream the select kernel is launched on, give
    the select kernel's mean time.  Returns (ms per step, mean select-kernel
    ms, last stats)."""
    z, cc, ty, tri, ft = outs
    stream = torch.cuda.current_stream().cuda_stream
    args = (d_cl.data_ptr(), n, mode, z.data_ptr(), cc.data_ptr(), ty.data_ptr(), tri.data_ptr(), ft.data_ptr())
    for _ in range(warmup):
        mesh.drop_device(spec, *args, stream=stream)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        mesh.drop_device(spec, *args, stream=stream)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    sel = []
    st = None
    for _ in range(steps):
        st = mesh.drop_device(spec, *args, stream=stream, stats=True)
        sel.append(st["select_ms"])
    return 1e3 * dt / steps, float(np.mean(sel)), st


def cpu_baseline(tris, spec, cl, n_kd, n_scan, seed, O, W):
    """The oracle on the host cores: reference KD-tree search (search_mode 0,
    the reference's algorithm) on a bounded sample, and the exhaustive overlap
    scan (search_mode 1) on a larger sample for the accuracy check."""
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    om = O.Mesh(tris)
    cut = O.cutter_from_spec(spec)
    idx_kd = np.sort(np.random.default_rng(seed).choice(len(cl), n_kd, replace=False))
    t0 = time.perf_counter()
    r_kd = om.batch_drop(cut, cl[idx_kd], search_mode=0, nthreads=threads)
    t_kd = time.perf_counter() - t0
    idx = np.sort(np.random.default_rng(seed + 1).choice(len(cl), n_scan, replace=False))
    t0 = time.perf_counter()
    r_scan = om.batch_drop(cut, cl[idx], search_mode=1, nthreads=threads)
    t_scan = time.perf_counter() - t0
    return dict(threads=threads, idx_kd=idx_kd, r_kd=r_kd, t_kd=t_kd, idx=idx, r=r_scan, t_scan=t_scan)


def traffic_from_profile(cutter):
    """HBM bytes per launch of the selection kernel, from the committed
    rocprofv3 FETCH_SIZE/WRITE_SIZE summary (copy-calibrated), or None."""
    path = os.path.join(ROOT, "profiles", f"drop_select_{cutter}_c1_hbm_bytes.md")
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("TRAFFIC_BYTES_PER_LAUNCH"):
            return float(line.split()[-1])
    return None


def profile_line(name, key, kernel):
    """The value of a `KEY kernel value` line of a committed rocprofv3
    summary under profiles/, or None."""
    path = os.path.join(ROOT, "profiles", name)
    if os.path.exists(path):
        for line in open(path):
            if line.startswith(f"{key} {kernel} "):
                return float(line.split()[-1])
    return None


def push_traffic(cutter):
    """HBM bytes per launch of the push kernel (mixed precision, X and Y
    fibers of the push workload), from the committed rocprofv3 summary."""
    return profile_line("push_fibers_c2_counters.md", "TRAFFIC_BYTES_PER_LAUNCH", f"k_push_fibers<{cutter}>")


def valu_from_profile(cutter):
    """VALU busy fraction and executed FP32 TFLOP/s of the selection kernel,
    from the committed rocprofv3 SQ counter summary, or None."""
    path = os.path.join(ROOT, "profiles", f"drop_select_{cutter}_c1_valu_counters.md")
    got = {}
    if os.path.exists(path):
        for line in open(path):
            if line.startswith(("VALU_BUSY_FRACTION", "EXECUTED_FP32_TFLOPS")):
                got[line.split()[0]] = float(line.split()[-1])
    return got or None


def overlap_pairs(tris, cl, R):
    """Number of (CL point, triangle) pairs that pass the reference's
    overlaps() test (millingcutter.cpp:422-433: the cutter's XY square
    against the triangle's XY box), for a CL set that is a full product
    grid: per triangle, the grid columns with x - R <= max x and x + R >=
    min x times the rows likewise, in the same FP64 arithmetic.  This is
    the pair count the reference's per-pair contact code runs on (before
    its below() test); culling lets the selection kernel skip most of it."""
    xs = np.unique(cl[:, 0])
    ys = np.unique(cl[:, 1])
    if len(xs) * len(ys) != len(cl):
        return None
    t = tris.reshape(-1, 3, 3)
    mnx, mxx = t[:, :, 0].min(axis=1), t[:, :, 0].max(axis=1)
    mny, mxy = t[:, :, 1].min(axis=1), t[:, :, 1].max(axis=1)
    cx = np.searchsorted(xs - R, mxx, side="right") - np.searchsorted(xs + R, mnx, side="left")
    cy = np.searchsorted(ys - R, mxy, side="right") - np.searchsorted(ys + R, mny, side="left")
    return int(np.sum(np.maximum(cx, 0).astype(np.int64) * np.maximum(cy, 0).astype(np.int64)))


def accuracy(got, ref, idx):
    dz = np.abs(got["z"][idx] - ref["z"])
    return {"max_dz_mm": float(dz.max()), "p99_dz_mm": float(np.percentile(dz, 99)),
            "type_mismatch": int(np.sum(got["type"][idx] != ref["type"])), "sample": int(len(idx))}


def cuda_ms(torch, fn, reps):
    """Mean device time of fn() over reps calls, HIP events on torch's
    current stream (where fn launches)."""
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        fn()
    e1.record()
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) / reps


def bench_backward(torch, capi, W, mesh, ntri, n, steps):
    """Differentiable drop on the 524k-triangle mesh: 1M uniform-random CL
    points, BallCutter d=10, FP32 forward (mixed) + FP32 backward with FP64
    edges, upstream gradient N(0,1)."""
    spec = SPECS["ball"]
    cl = W.cl_random(n, 200.0, z0=-10.0, seed=44)
    d_cl = torch.from_numpy(cl).cuda()
    z = torch.empty(n, dtype=torch.float64, device="cuda")
    cc = torch.empty((n, 3), dtype=torch.float64, device="cuda")
    ty = torch.empty(n, dtype=torch.uint8, device="cuda")
    tri = torch.empty(n, dtype=torch.int32, device="cuda")
    ft = torch.empty(n, dtype=torch.uint8, device="cuda")
    gz = torch.randn(n, dtype=torch.float64, device="cuda", generator=torch.Generator("cuda").manual_seed(4))
    dz = torch.empty((n, 3), dtype=torch.float64, device="cuda")
    gtri = torch.zeros((ntri, 9), dtype=torch.float64, device="cuda")
    gr = torch.zeros(1, dtype=torch.float64, device="cuda")
    stream = torch.cuda.current_stream().cuda_stream

    def fwd():
        mesh.drop_device(spec, d_cl.data_ptr(), n, "mixed", z.data_ptr(), cc.data_ptr(), ty.data_ptr(),
                         tri.data_ptr(), ft.data_ptr(), stream=stream)

    def bwd():
        mesh.drop_backward(spec, d_cl.data_ptr(), n, cc.data_ptr(), ty.data_ptr(), tri.data_ptr(), ft.data_ptr(),
                           d_grad_z=gz.data_ptr(), precision="mixed", d_dz=dz.data_ptr(),
                           d_grad_tri=gtri.data_ptr(), d_grad_r=gr.data_ptr(), stream=stream)

    fwd()
    bwd()
    torch.cuda.synchronize()
    f_ms = cuda_ms(torch, fwd, steps)
    b_ms = cuda_ms(torch, bwd, steps)
    # algorithmic HBM bytes per point of the backward kernel: CL xyz, CC xyz,
    # type/tri/feat, upstream, 9 vertex doubles of the contact triangle, dz/dx,
    # dz/dy, dz/dr, and the 9 read-modify-write FP64 vertex sums
    bpp = 24 + 24 + 6 + 8 + 72 + 24 + 9 * 16
    return {"workload": f"524,288-triangle heightfield, BallCutter d=10, {n} uniform-random CL points, "
                        "upstream gradient N(0,1)",
            "forward_ms": f_ms, "backward_ms": b_ms, "value": n / ((f_ms + b_ms) / 1e3),
            "unit": "CL pts/s (forward + backward)",
            "backward_roofline": {"bound": "hbm", "bytes_per_point": bpp,
                                  "achieved": n * bpp / (b_ms * 1e-3) / 1e9, "peak": 8000.0, "unit": "GB/s",
                                  "frac": n * bpp / (b_ms * 1e-3) / 1e9 / 8000.0}}


PUSH_SPECS = {"bull": {"kind": "bull", "d": 10.0, "r": 2.0, "L": 20.0},
              "cone": {"kind": "cone", "d": 10.0, "angle": np.pi / 4, "L": 10.0}}


def bench_push(capi, W, cpu_fibers, do_cpu):
    """Waterline fibers on the 204,802-triangle sphere scene: 2,000 X and
    2,000 Y fibers at each of 20 levels (80,000 fibers) per cutter, the
    reference's KD-search candidate set, FP64, X and Y in one launch
    (ocl_hip_push_xy).  A first call builds both directions' trees and sizes
    the interval pools; the second is timed.  Kernel time from HIP events on
    the launch stream (ocl_stats.kernel_ms)."""
    tris = W.config2_mesh()
    levels = W.config2_levels()
    dev = capi.DeviceMesh(tris)
    out = {"workload": f"{len(tris)}-triangle sphere scene, 20 levels x (2000 X + 2000 Y fibers), one launch"}
    om = None
    for name, spec in PUSH_SPECS.items():
        fibs = {"x": [], "y": []}
        for z in levels:
            xf, yf = W.waterline_fibers(tris, spec["d"] / 2, z, 2000)
            fibs["x"].append(xf)
            fibs["y"].append(yf)
        fibs = {d: np.concatenate(v) for d, v in fibs.items()}
        # every candidate pair, as the reference pushes them (EXACT64)
        dev.push_xy(spec, fibs["x"], fibs["y"], precision="exact64")  # warm-up: trees and pool sizes
        ra = dev.push_xy(spec, fibs["x"], fibs["y"], precision="exact64")
        sa = ra["stats"]
        # the product default (MIXED): pairs that cannot change their fiber
        # left out, same intervals
        dev.push_xy(spec, fibs["x"], fibs["y"], precision="mixed")
        t0 = time.perf_counter()
        r = dev.push_xy(spec, fibs["x"], fibs["y"], precision="mixed")
        wall = time.perf_counter() - t0
        st = r["stats"]
        kms = st["kernel_ms"]
        same = all(np.array_equal(r[d]["offsets"], ra[d]["offsets"]) and np.array_equal(r[d]["lohi"], ra[d]["lohi"])
                   and np.array_equal(r[d]["cc"], ra[d]["cc"]) for d in ("x", "y"))
        nint = len(r["x"]["lohi"]) + len(r["y"]["lohi"])
        nf = len(fibs["x"]) + len(fibs["y"])
        ach = st["pairs"] * F_PUSH_KEPT[name] / (kms * 1e-3) / 1e12
        ref_tf = sa["pairs"] * F_PUSH[name] / (kms * 1e-3) / 1e12
        line = {"fibers": nf, "intervals": nint, "pairs": int(st["pairs"]), "pairs_skipped": int(st["skipped"]),
                "brent_iters": int(st["brent_iters"]),
                "kernel_ms": kms, "fibers_per_s": nf / (kms / 1e3), "intervals_per_s": nint / (kms / 1e3),
                "wall_s_incl_host": wall,
                "every_pair": {"kernel_ms": sa["kernel_ms"], "pairs": int(sa["pairs"]),
                               "brent_iters": int(sa["brent_iters"]),
                               "fibers_per_s": nf / (sa["kernel_ms"] / 1e3), "same_intervals": same},
                "speedup_vs_every_pair": sa["kernel_ms"] / kms,
                "roofline": {"bound": "valu", "kernel": f"k_push_fibers<{name}>", "achieved": ach,
                             "peak": PEAK_FP64, "unit": "TFLOP/s (FP64)", "frac": ach / PEAK_FP64,
                             "flop_per_pair": F_PUSH_KEPT[name], "pairs_per_launch": int(st["pairs"]),
                             "kernel_ms": kms, "traffic": push_traffic(name),
                             "traffic_source": "profiles/push_fibers_c2_counters.md",
                             "counters": {"valu_busy": profile_line("push_fibers_c2_counters.md", "VALU_BUSY_FRACTION",
                                                                    f"k_push_fibers<{name}>"),
                                          "executed_fp64_tflops": profile_line(
                                              "push_fibers_c2_counters.md", "EXECUTED_FP64_TFLOPS",
                                              f"k_push_fibers<{name}>")},
                             "reference_equivalent": {"pairs": int(sa["pairs"]), "flop_per_pair": F_PUSH[name],
                                                      "tflops": ref_tf, "frac": ref_tf / PEAK_FP64}}}
        if do_cpu:
            from oracle import oracle as O
            if om is None:
                om = O.Mesh(tris)
            threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
            idx = np.sort(np.random.default_rng(7).choice(len(fibs["x"]), cpu_fibers, replace=False))
            t0 = time.perf_counter()
            om.batch_push(O.cutter_from_spec(spec), fibs["x"][idx], 0, search_mode=0, nthreads=threads)
            tc = time.perf_counter() - t0
            line["cpu_baseline"] = {"value": cpu_fibers / tc, "unit": "fibers/s", "cores": threads, "kind": "port",
                                    "sample": f"{cpu_fibers} seeded X fibers of the same set, FP64 oracle with "
                                              "the reference's KD-search candidates"}
        out[name] = line
        log(f"push {name}: {nf} fibers, {nint} intervals, kernel {kms:.1f} ms (every pair {sa['kernel_ms']:.1f} ms), "
            f"{st['pairs']} pairs + {st['skipped']} skipped, same intervals {same}")
    dev.close()
    return out


def bench_refine(W, tris, tol=0.01, sampling=1.0, min_step=0.01, z0=-20.0):
    """Slope-driven refinement demo: 500 random polylines (3-6 vertices,
    seed 44, over 200 x 200 mm) on the 524,288-triangle heightfield,
    BallCutter d=10 L=20, ocl::AdaptivePathDropCutter to a 0.01 mm chord
    tolerance (first samples 1 mm apart, minimum step 0.01 mm): the C++ host
    loop refines level by level, each level one device batch (mixed
    precision) of the midpoints and quarter points still open.  Times run()
    (setSTL, the device mesh build, is outside)."""
    from opencamlib_amd import ocl
    polys = W.random_polylines(500, 200.0, seed=44)
    surf = ocl.STLSurf()
    for t in np.asarray(tris).reshape(-1, 9):
        surf.addTriangle(ocl.Triangle(ocl.Point(*t[0:3]), ocl.Point(*t[3:6]), ocl.Point(*t[6:9])))
    path = ocl.Path()
    for poly in polys:
        for a, b in zip(poly[:-1], poly[1:]):
            path.append(ocl.Line(ocl.Point(a[0], a[1], z0), ocl.Point(b[0], b[1], z0)))
    apdc = ocl.AdaptivePathDropCutter()
    apdc.setSTL(surf)
    apdc.setCutter(ocl.BallCutter(10.0, 20.0))
    apdc.setPath(path)
    apdc.setSampling(sampling)
    apdc.setMinSampling(min_step)
    apdc.setChordTolerance(tol)
    t0 = time.perf_counter()
    apdc.run()
    secs = time.perf_counter() - t0
    npts = len(apdc.getCLPoints())
    length = float(sum(np.sum(np.hypot(*np.diff(p, axis=0).T)) for p in polys))
    return {"workload": "500 random polylines (164 m of path) on the 524,288-triangle heightfield, BallCutter "
                        "d=10, AdaptivePathDropCutter to 0.01 mm chord tolerance (1 mm first samples, 0.01 mm "
                        "minimum step)",
            "path_mm": length, "points": npts, "levels": apdc.getLevels(), "seconds": secs,
            "value": npts / secs, "unit": "refined CL pts/s (host loop + device batches)",
            "uniform_points_at_min_step": int(length / min_step),
            "points_saved_vs_uniform": length / min_step / npts}


def bench_batch(torch, capi, W, steps):
    """64 uneven jobs (2k-500k-triangle heightfields, all five cutter kinds,
    1k-1M CL points each) in mixed precision: autotuned persistent launch in
    a hipGraph vs one selection launch per job vs one ocl_hip_drop_device
    call per job."""
    t0 = time.perf_counter()
    jobs = W.config3_jobs(64)
    meshes = [capi.DeviceMesh(j["tris"]) for j in jobs]
    t_build = time.perf_counter() - t0
    cls = [torch.from_numpy(j["cl"]).cuda() for j in jobs]
    outs = []
    for j in jobs:
        n = len(j["cl"])
        outs.append([torch.empty(n, dtype=torch.float64, device="cuda"),
                     torch.empty((n, 3), dtype=torch.float64, device="cuda"),
                     torch.empty(n, dtype=torch.uint8, device="cuda"),
                     torch.empty(n, dtype=torch.int32, device="cuda"),
                     torch.empty(n, dtype=torch.uint8, device="cuda")])
    items = [{"mesh": m, "spec": j["spec"], "cl": c.data_ptr(), "n": len(j["cl"]),
              "z": o[0].data_ptr(), "cc": o[1].data_ptr(), "type": o[2].data_ptr(), "tri": o[3].data_ptr(),
              "feat": o[4].data_ptr()} for m, j, c, o in zip(meshes, jobs, cls, outs)]
    cj = capi.make_jobs(items)
    total = sum(len(j["cl"]) for j in jobs)
    t0 = time.perf_counter()
    best, best_ms = capi.autotune(cj, "mixed", trials=2)
    t_tune = time.perf_counter() - t0
    log(f"batch: {len(jobs)} jobs, {sum(len(j['tris']) for j in jobs)} triangles, {total} CL points; "
        f"meshes {t_build:.1f} s, autotune {t_tune:.1f} s -> {best} ({best_ms:.2f} ms)")
    stream = torch.cuda.current_stream().cuda_stream

    def timed(batch):
        batch.run(stream=stream, want_stats=False)
        torch.cuda.synchronize()
        t = time.perf_counter()
        for _ in range(steps):
            batch.run(stream=stream, want_stats=False)
        torch.cuda.synchronize()
        return 1e3 * (time.perf_counter() - t) / steps

    b = capi.Batch(cj, "mixed", {**best, "use_graph": 1})
    ms_graph = timed(b)
    b.close()
    # the persistent selection launch alone (eager run, HIP events around it)
    # and each job's evaluated pairs, for the selection kernel's roofline
    b = capi.Batch(cj, "mixed", {**best, "use_graph": 0})
    b.run(stream=stream, want_stats=False)
    st = b.run(stream=stream, want_stats=True)
    sel_flops = 0.0
    sel_pairs = 0
    for jn, j in enumerate(jobs):
        if j["spec"]["kind"] in F_PAIR:
            p = b.job_stats(jn)["pairs"]
            sel_pairs += p
            sel_flops += p * F_PAIR[j["spec"]["kind"]]
    b.close()
    sel_ach = sel_flops / (st["select_ms"] * 1e-3) / 1e12
    b = capi.Batch(cj, "mixed", {"per_job_launch": 1})
    ms_perjob = timed(b)
    b.close()

    def singles():
        for m, j, c, o in zip(meshes, jobs, cls, outs):
            m.drop_device(j["spec"], c.data_ptr(), len(j["cl"]), "mixed", *(t.data_ptr() for t in o),
                          stream=stream)

    singles()
    torch.cuda.synchronize()
    t = time.perf_counter()
    for _ in range(steps):
        singles()
    torch.cuda.synchronize()
    ms_single = 1e3 * (time.perf_counter() - t) / steps
    for m in meshes:
        m.close()
    kinds = {}
    for j in jobs:
        kinds[j["spec"]["kind"]] = kinds.get(j["spec"]["kind"], 0) + 1
    return {"workload": f"{len(jobs)} jobs, {sum(len(j['tris']) for j in jobs)} triangles, {total} CL points, "
                        f"cutters {kinds}",
            "tune": best, "autotune_s": t_tune,
            "graph_persistent": {"ms_per_step": ms_graph, "value": total / (ms_graph / 1e3)},
            "per_job_launch": {"ms_per_step": ms_perjob, "value": total / (ms_perjob / 1e3)},
            "call_per_job": {"ms_per_step": ms_single, "value": total / (ms_single / 1e3)},
            "unit": "CL pts/s", "speedup_vs_call_per_job": ms_single / ms_graph,
            "roofline": {"bound": "valu", "kernel": "k_drop_select_jobs<float>", "achieved": sel_ach,
                         "peak": PEAK_FP32, "unit": "TFLOP/s", "frac": sel_ach / PEAK_FP32,
                         "frac_unpacked": sel_ach / PEAK_FP32_UNPACKED, "pairs_per_launch": int(sel_pairs),
                         "flop_per_pair": "per job: " + json.dumps(F_PAIR), "kernel_ms": st["select_ms"],
                         "eager_ms_per_step": st["kernel_ms"],
                         "traffic": profile_line("batch_c3_counters.md", "TRAFFIC_BYTES_PER_LAUNCH",
                                                 "k_drop_select_jobs<float>"),
                         "traffic_source": "profiles/batch_c3_counters.md",
                         "counters": {"valu_busy": profile_line("batch_c3_counters.md", "VALU_BUSY_FRACTION",
                                                                "k_drop_select_jobs<float>"),
                                      "executed_fp32_tflops": profile_line("batch_c3_counters.md",
                                                                           "EXECUTED_FP32_TFLOPS",
                                                                           "k_drop_select_jobs<float>")}}}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--cutter", default="ball", choices=sorted(SPECS))
    ap.add_argument("--side", type=int, default=1000, help="CL grid side (1000 = 1M points)")
    ap.add_argument("--cpu-kd-points", type=int, default=100_000)
    ap.add_argument("--cpu-scan-points", type=int, default=100_000)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-compare", action="store_true", help="skip the FP64/FP32/exact device lines")
    ap.add_argument("--no-extra", action="store_true",
                    help="skip the backward, push and batched-jobs sections")
    ap.add_argument("--cpu-push-fibers", type=int, default=200)
    ap.add_argument("--only", choices=("push", "batch", "backward"),
                    help="run one extra section alone and print its JSON object (for profiling)")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="first run 3 device copies of 256 MiB (reference for HBM byte counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; multi-GPU sharding is a design only",
                          "n_gpus": args.gpus}))
        sys.exit(2)

    import torch
    from opencamlib_amd import capi, workloads as W

    if not torch.cuda.is_available():
        sys.exit("bench.py needs a HIP device")
    torch.cuda.set_device(0)
    if args.calibrate_copy:
        src = torch.ones(256 << 18, dtype=torch.float32, device="cuda")  # 256 MiB
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        log("copy calibration: 3 x 268435456 bytes read + written")
        del src, dst
    if args.only == "push":
        print(json.dumps({"push": bench_push(capi, W, args.cpu_push_fibers, not args.no_cpu)}), flush=True)
        return
    if args.only == "batch":
        print(json.dumps({"batch": bench_batch(torch, capi, W, max(3, args.steps // 2))}), flush=True)
        return
    if args.only == "backward":
        tris = W.config1_mesh()
        mesh = capi.DeviceMesh(tris)
        print(json.dumps({"backward": bench_backward(torch, capi, W, mesh, len(tris), 1_000_000,
                                                     max(3, args.steps // 2))}), flush=True)
        mesh.close()
        return
    spec = SPECS[args.cutter]
    t0 = time.perf_counter()
    tris = W.config1_mesh()
    cl = W.config1_cl(args.side)
    n = len(cl)
    log(f"inputs: {len(tris)} triangles, {n} CL points ({time.perf_counter() - t0:.1f} s)")
    t0 = time.perf_counter()
    mesh = capi.DeviceMesh(tris)
    log(f"device mesh built in {time.perf_counter() - t0:.2f} s")

    d_cl = torch.from_numpy(cl).cuda()
    outs = (torch.empty(n, dtype=torch.float64, device="cuda"),
            torch.empty((n, 3), dtype=torch.float64, device="cuda"),
            torch.empty(n, dtype=torch.uint8, device="cuda"),
            torch.empty(n, dtype=torch.int32, device="cuda"),
            torch.empty(n, dtype=torch.uint8, device="cuda"))

    def fetch():
        z, cc, ty, tri, ft = outs
        return {"z": z.cpu().numpy().copy(), "type": ty.cpu().numpy().copy()}

    ms, sel_ms, st = time_mode(mesh, spec, d_cl, n, "mixed", outs, args.steps, args.warmup, torch)
    res_mixed = fetch()
    log(f"mixed: {ms:.2f} ms/step, select kernel {sel_ms:.2f} ms, stats {st}")
    lines = {}
    results = {}
    if not args.no_compare:
        for mode, k in (("fp64", max(3, args.steps // 2)), ("fp32", max(3, args.steps // 2)), ("exact64", 2)):
            m_ms, m_sel, m_st = time_mode(mesh, spec, d_cl, n, mode, outs, k, 1, torch)
            results[mode] = fetch()
            lines[mode] = {"value": n / (m_ms / 1e3), "ms_per_step": m_ms, "select_ms": m_sel,
                           "escalated": m_st["escalated"]}
            log(f"{mode}: {m_ms:.2f} ms/step, select kernel {m_sel:.2f} ms")

    pairs = st["pairs"]
    flops = pairs * F_PAIR[args.cutter]
    achieved = flops / (sel_ms * 1e-3) / 1e12
    roofline = {"bound": "valu", "kernel": f"k_drop_select<float,{args.cutter}>",
                "achieved": achieved, "peak": PEAK_FP32, "unit": "TFLOP/s", "frac": achieved / PEAK_FP32,
                "peak_unpacked_fp32": PEAK_FP32_UNPACKED, "frac_unpacked": achieved / PEAK_FP32_UNPACKED,
                "pairs_per_launch": int(pairs), "flop_per_pair": F_PAIR[args.cutter],
                "wave_evals": int(st["wave_evals"]),
                "lane_utilisation": pairs / (64.0 * st["wave_evals"]) if st["wave_evals"] else None,
                "kernel_ms": sel_ms, "traffic": traffic_from_profile(args.cutter),
                "traffic_source": f"profiles/drop_select_{args.cutter}_c1_hbm_bytes.md"}
    n_overlap = overlap_pairs(tris, cl, spec["d"] / 2.0)
    if n_overlap is not None:
        # the reference's work on the same input: its contact code on every
        # overlapping pair, at this line's whole-step time (not `achieved`)
        ref_tf = n_overlap * F_PAIR[args.cutter] / (ms * 1e-3) / 1e12
        roofline["reference_equivalent"] = {
            "overlap_pairs": n_overlap, "evaluated_fraction": pairs / n_overlap, "tflops": ref_tf,
            "frac": ref_tf / PEAK_FP32, "frac_unpacked": ref_tf / PEAK_FP32_UNPACKED}
    counters = valu_from_profile(args.cutter)
    if counters:
        roofline["counters"] = {"valu_busy": counters.get("VALU_BUSY_FRACTION"),
                                "executed_fp32_tflops": counters.get("EXECUTED_FP32_TFLOPS"),
                                "source": f"profiles/drop_select_{args.cutter}_c1_valu_counters.md"}

    out = {"metric": METRIC, "value": n / (ms / 1e3), "unit": "CL pts/s", "n_gpus": 1, "steps": args.steps,
           "warmup": args.warmup, "ms_per_step": ms, "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "f32+f64", "data": "synthetic (seeded heightfield, regular CL grid)",
           "config": {"workload": f"batch drop-cutter, {args.cutter} d=10 mm, 524,288-triangle heightfield with "
                                  f"8 bumps + 20 steps, {n} CL points (grid 0.2 mm)",
                      "triangles": len(tris), "cl_points": n, "cutter": spec, "precision": "mixed"},
           "escalated": {"first_pass": st["escalated"], "to_exact": st["escalated_exact"]},
           "roofline": roofline}
    if lines:
        out["fp64_gpu"] = lines["fp64"]
        out["fp32_gpu"] = lines["fp32"]
        out["exact64_gpu"] = lines["exact64"]
        out["mixed_over_fp64"] = out["value"] / lines["fp64"]["value"]

    if not args.no_cpu:
        from oracle import oracle as O
        t0 = time.perf_counter()
        cb = cpu_baseline(tris, spec, cl, args.cpu_kd_points, args.cpu_scan_points, 99, O, W)
        log(f"oracle: KD {args.cpu_kd_points} pts in {cb['t_kd']:.1f} s, scan {args.cpu_scan_points} pts in "
            f"{cb['t_scan']:.1f} s ({time.perf_counter() - t0:.1f} s incl. build)")
        out["cpu_baseline"] = {
            "value": args.cpu_kd_points / cb["t_kd"], "unit": "CL pts/s", "cores": cb["threads"], "kind": "port",
            "sample": f"{args.cpu_kd_points} seeded CL points of the same 1M grid, FP64 oracle with the "
                      f"reference's KD-tree search (bucket 1), OpenMP dynamic over points, {cb['threads']} threads",
            "host": host_cpu(),
            "note": "a C restatement of the reference's algorithm (same KD-tree split and search, same contact "
                    "math), not the reference: it skips the reference's per-query std::list copy of every "
                    "candidate triangle (kdtree.hpp:276-277), which dominates the reference's time",
            "reference_measured": REFERENCE_MEASURED}
        out["cpu_scan_oracle"] = {"value": args.cpu_scan_points / cb["t_scan"], "unit": "CL pts/s",
                                  "cores": cb["threads"],
                                  "sample": f"{args.cpu_scan_points} seeded CL points, exhaustive overlap scan "
                                            f"in KD order (same results as the KD search)"}
        acc = {"mixed": accuracy(res_mixed, cb["r"], cb["idx"])}
        for mode, r in results.items():
            acc[mode] = accuracy(r, cb["r"], cb["idx"])
        acc["mixed_vs_kd_sample"] = accuracy(res_mixed, cb["r_kd"], cb["idx_kd"])
        out["accuracy_vs_fp64_oracle"] = acc
    if not args.no_extra and args.cutter == "ball":
        # the config's second run: BullCutter d=10 r=2 on the same mesh and grid
        bspec = SPECS["bull"]
        b_ms, b_sel, b_st = time_mode(mesh, bspec, d_cl, n, "mixed", outs, max(3, args.steps // 2), 1, torch)
        b_mixed = fetch()
        f_ms, f_sel, _ = time_mode(mesh, bspec, d_cl, n, "fp64", outs, 3, 1, torch)
        b_ach = b_st["pairs"] * F_PAIR["bull"] / (b_sel * 1e-3) / 1e12
        bull = {"mixed": {"value": n / (b_ms / 1e3), "ms_per_step": b_ms, "select_ms": b_sel,
                          "escalated": b_st["escalated"], "brent_iters": b_st["brent_iters"]},
                "fp64_gpu": {"value": n / (f_ms / 1e3), "ms_per_step": f_ms, "select_ms": f_sel},
                "mixed_over_fp64": f_ms / b_ms,
                "roofline": {"bound": "valu", "kernel": "k_drop_select<float,bull>", "achieved": b_ach,
                             "peak": PEAK_FP32, "unit": "TFLOP/s", "frac": b_ach / PEAK_FP32,
                             "frac_unpacked": b_ach / PEAK_FP32_UNPACKED, "pairs_per_launch": int(b_st["pairs"]),
                             "flop_per_pair": F_PAIR["bull"], "kernel_ms": b_sel,
                             "lane_utilisation": (b_st["pairs"] / (64.0 * b_st["wave_evals"])
                                                  if b_st["wave_evals"] else None),
                             "traffic": traffic_from_profile("bull"),
                             "traffic_source": "profiles/drop_select_bull_c1_hbm_bytes.md"}}
        bc = valu_from_profile("bull")
        if bc:
            bull["roofline"]["counters"] = {"valu_busy": bc.get("VALU_BUSY_FRACTION"),
                                            "executed_fp32_tflops": bc.get("EXECUTED_FP32_TFLOPS"),
                                            "source": "profiles/drop_select_bull_c1_valu_counters.md"}
        if not args.no_cpu:
            from oracle import oracle as O
            idx = np.sort(np.random.default_rng(100).choice(n, 20_000, replace=False))
            rb = O.Mesh(tris).batch_drop(O.cutter_from_spec(bspec), cl[idx], search_mode=1)
            bull["accuracy_vs_fp64_oracle"] = accuracy(b_mixed, rb, idx)
        out["bull"] = bull
        log(f"bull: mixed {b_ms:.2f} ms, fp64 {f_ms:.2f} ms")
    if not args.no_extra:
        out["backward"] = bench_backward(torch, capi, W, mesh, len(tris), 1_000_000, max(3, args.steps // 2))
        log(f"backward: forward {out['backward']['forward_ms']:.2f} ms, backward {out['backward']['backward_ms']:.3f} ms")
    mesh.close()
    if not args.no_extra:
        out["push"] = bench_push(capi, W, args.cpu_push_fibers, not args.no_cpu)
        out["batch"] = bench_batch(torch, capi, W, max(3, args.steps // 2))
        out["refine"] = bench_refine(W, tris)
        log(f"refine: {out['refine']['points']} points in {out['refine']['seconds']:.2f} s, "
            f"{out['refine']['levels']} levels")
        log(f"batch: graph {out['batch']['graph_persistent']['ms_per_step']:.2f} ms, per-job launch "
            f"{out['batch']['per_job_launch']['ms_per_step']:.2f} ms, call per job "
            f"{out['batch']['call_per_job']['ms_per_step']:.2f} ms")
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
