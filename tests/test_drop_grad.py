"""Drop-cutter gradients: dz/dx, dz/dy, dz/dr and dz/d(triangle vertices).

CPU: the oracle's FP64 analytic gradient (oc_drop_grad) against central
finite differences of the oracle's own forward drop (h = 1e-4 mm), at CL
points at least 0.01 mm from a contact switch (the active (triangle,
feature) is the same at +-0.01 mm in x and y).  This pins the analytic
formulas: a wrong sign, a dropped term or a swapped weight fails it.

GPU: the device backward kernel (through the C ABI) against the oracle's
analytic gradient on the same forward contacts; FP32 arithmetic with the
edge contacts in FP64, relative tolerance 2e-4 (FP64 mode: 1e-9).
"""
import math

import numpy as np
import pytest

from cutters import BALL10, BULL10, CYL6, CYLBALL
from opencamlib_amd import workloads as W
from oracle import oracle as O

H = 1e-4        # central-difference step (mm)
SWITCH = 0.01   # distance from a contact switch (mm)
# cone with l = 0: the reference's edge formula then is the true cone contact
# (conecutter.cpp:115-122 use length for center_height)
CONE_L0 = {"kind": "cone", "d": 10.0, "angle": math.pi / 4, "L": 0.0}
# with a shank (l = 10): the reference's edge formula is then not the cone's
# envelope, and the gradient differentiates the formula as written
CONE_L10 = {"kind": "cone", "d": 10.0, "angle": math.pi / 4, "L": 10.0}
CASES = {"cyl": CYL6, "ball": BALL10, "bull": BULL10, "cone": CONE_L0, "cone_shank": CONE_L10, "cylball": CYLBALL}


def _spec_d(spec, dd):
    s = dict(spec)
    s["d"] = spec["d"] + dd
    return s


def _stable_points(mesh, cut, cl, res):
    """Points whose active contact is the same at +-SWITCH in x and y."""
    ok = res["tri"] >= 0
    for ax in (0, 1):
        for sgn in (-1, 1):
            c2 = cl.copy()
            c2[:, ax] += sgn * SWITCH
            r2 = mesh.batch_drop(cut, c2, search_mode=1)
            ok &= (r2["tri"] == res["tri"]) & (r2["feat"] == res["feat"])
    return ok


@pytest.fixture(scope="module")
def grad_scene():
    tris = W.heightfield(32, 40.0, bumps=3, steps=2, seed=5)
    cl = W.cl_random(1500, 40.0, z0=-10.0, seed=8)
    return tris, cl


@pytest.mark.parametrize("name", list(CASES))
def test_analytic_gradient_matches_central_differences(grad_scene, name):
    tris, cl = grad_scene
    spec = CASES[name]
    mesh = O.Mesh(tris)
    cut = O.cutter_from_spec(spec)
    res = mesh.batch_drop(cut, cl, search_mode=1)
    g = mesh.batch_drop_grad(cut, cl, res)
    ok = _stable_points(mesh, cut, cl, res)
    idx = np.nonzero(ok)[0]
    assert len(idx) > 300
    # dz/dx, dz/dy on the full mesh
    fd = np.zeros((len(idx), 3))
    for ax in (0, 1):
        cp, cm = cl[idx].copy(), cl[idx].copy()
        cp[:, ax] += H
        cm[:, ax] -= H
        zp = mesh.batch_drop(cut, cp, search_mode=1)["z"]
        zm = mesh.batch_drop(cut, cm, search_mode=1)["z"]
        fd[:, ax] = (zp - zm) / (2 * H)
    # dz/dr: the same contact with the diameter changed by +-2H (composite
    # cutters have no single radius: their dz/dr is NaN and not checked)
    if spec["kind"] != "composite":
        zp = mesh.batch_drop(O.cutter_from_spec(_spec_d(spec, 2 * H)), cl[idx], search_mode=1)["z"]
        zm = mesh.batch_drop(O.cutter_from_spec(_spec_d(spec, -2 * H)), cl[idx], search_mode=1)["z"]
        fd[:, 2] = (zp - zm) / (2 * H)
    else:
        assert np.all(np.isnan(g[idx, 2]))
        g = g.copy()
        g[:, 2] = 0.0
    err = np.abs(fd - g[idx, :3])
    tol = 2e-5 + 2e-5 * np.abs(fd)
    bad = np.nonzero(np.any(err > tol, axis=1))[0]
    print(f"[grad-fd] {name}: points={len(idx)} max|d(xy)|={err[:, :2].max():.2e} max|dr|={err[:, 2].max():.2e} "
          f"bad={len(bad)}", flush=True)
    # a point can sit within H of a switch in r even when stable in x, y:
    # allow a tiny fraction, never a systematic error
    assert len(bad) <= max(2, len(idx) // 200), (name, fd[bad[:5]], g[idx][bad[:5], :3])


@pytest.mark.parametrize("name", list(CASES))
def test_vertex_gradient_matches_central_differences(grad_scene, name):
    tris, cl = grad_scene
    spec = CASES[name]
    mesh = O.Mesh(tris)
    cut = O.cutter_from_spec(spec)
    res = mesh.batch_drop(cut, cl[:400], search_mode=1)
    g = mesh.batch_drop_grad(cut, cl[:400], res)
    ok = _stable_points(mesh, cut, cl[:400], res)
    checked = worst = 0
    seen_feats = set()
    for i in np.nonzero(ok)[0][:120]:
        t = tris[res["tri"][i]].copy()
        fdv = np.zeros(9)
        same = True
        for j in range(9):
            zs = []
            for sgn in (1, -1):
                tt = t.copy()
                tt[j] += sgn * H
                r = O.drop_triangle(cut, tt, cl[i, 0], cl[i, 1])
                same &= r["feat"] == res["feat"][i]
                zs.append(r["z"])
            fdv[j] = (zs[0] - zs[1]) / (2 * H)
        if not same:
            continue
        checked += 1
        seen_feats.add(res["feat"][i] & 15)
        e = np.max(np.abs(fdv - g[i, 3:]) / (1.0 + np.abs(fdv)))
        worst = max(worst, e)
        assert e < 5e-5, (name, i, res["feat"][i], fdv, g[i, 3:])
    print(f"[grad-fd-vertex] {name}: checked={checked} features={sorted(seen_feats)} worst={worst:.2e}", flush=True)
    assert checked > 40


def test_gradient_invariants():
    """Translating the CL point and the triangle together leaves z unchanged:
    dz/dx + sum of dz/dp.x = 0 (same for y), and lifting the triangle lifts z
    one for one: sum of dz/dp.z = 1."""
    tris = W.heightfield(16, 20.0, bumps=2, steps=1, seed=2)
    cl = W.cl_random(400, 20.0, seed=3)
    mesh = O.Mesh(tris)
    for spec in (CYL6, BALL10, BULL10, CONE_L0, CYLBALL):
        cut = O.cutter_from_spec(spec)
        res = mesh.batch_drop(cut, cl, search_mode=1)
        g = mesh.batch_drop_grad(cut, cl, res)
        m = res["tri"] >= 0
        gv = g[m, 3:].reshape(-1, 3, 3)
        assert np.allclose(g[m, 0] + gv[:, :, 0].sum(1), 0.0, atol=1e-9)
        assert np.allclose(g[m, 1] + gv[:, :, 1].sum(1), 0.0, atol=1e-9)
        assert np.allclose(gv[:, :, 2].sum(1), 1.0, atol=1e-9)


# ----------------------------------------------------------------- GPU

def _device_forward(dev, spec, cl, precision):
    import torch
    n = len(cl)
    d_cl = torch.tensor(cl, dtype=torch.float64, device="cuda")
    z = torch.empty(n, dtype=torch.float64, device="cuda")
    cc = torch.empty((n, 3), dtype=torch.float64, device="cuda")
    ty = torch.empty(n, dtype=torch.uint8, device="cuda")
    tri = torch.empty(n, dtype=torch.int32, device="cuda")
    ft = torch.empty(n, dtype=torch.uint8, device="cuda")
    dev.drop_device(spec, d_cl.data_ptr(), n, precision, z.data_ptr(), cc.data_ptr(), ty.data_ptr(),
                    tri.data_ptr(), ft.data_ptr())
    torch.cuda.synchronize()
    return d_cl, z, cc, ty, tri, ft


def _device_backward(dev, spec, fw, ntri, gz, precision):
    import torch
    d_cl, z, cc, ty, tri, ft = fw
    n = d_cl.shape[0]
    dz = torch.empty((n, 3), dtype=torch.float64, device="cuda")
    gtri = torch.zeros((ntri, 9), dtype=torch.float64, device="cuda")
    gr = torch.zeros(1, dtype=torch.float64, device="cuda")
    g = torch.tensor(gz, dtype=torch.float64, device="cuda")
    dev.drop_backward(spec, d_cl.data_ptr(), n, cc.data_ptr(), ty.data_ptr(), tri.data_ptr(), ft.data_ptr(),
                      d_grad_z=g.data_ptr(), precision=precision, d_dz=dz.data_ptr(), d_grad_tri=gtri.data_ptr(),
                      d_grad_r=gr.data_ptr())
    torch.cuda.synchronize()
    return dz.cpu().numpy(), gtri.cpu().numpy(), float(gr.item())


def _oracle_grads(mesh, spec, cl, fw):
    _, z, cc, ty, tri, ft = fw
    res = dict(tri=tri.cpu().numpy().astype(np.int32), feat=ft.cpu().numpy().astype(np.int32),
               type=ty.cpu().numpy().astype(np.int32), cc=cc.cpu().numpy())
    return mesh.batch_drop_grad(O.cutter_from_spec(spec), cl, res), res


def _check_device_vs_oracle(tris, spec, cl, fwd_prec, bwd_prec, label):
    from opencamlib_amd import capi
    rng = np.random.default_rng(4)
    gz = rng.normal(size=len(cl))
    dev = capi.DeviceMesh(tris)
    fw = _device_forward(dev, spec, cl, fwd_prec)
    dz, gtri, gr = _device_backward(dev, spec, fw, len(tris), gz, bwd_prec)
    dev.close()
    mesh = O.Mesh(tris)
    g, res = _oracle_grads(mesh, spec, cl, fw)
    has = res["tri"] >= 0
    if bwd_prec in ("fp64", "exact64"):
        rtol, atol = 1e-9, 1e-9
    else:
        rtol, atol = 2e-4, 2e-5
    ref_xyr = g[:, :3].copy()
    got_xyr = dz.copy()
    if spec["kind"] == "composite":
        assert np.all(np.isnan(got_xyr[has, 2]))
        ref_xyr[:, 2] = got_xyr[:, 2] = 0.0
    err = np.abs(got_xyr - ref_xyr)
    tol = atol + rtol * np.abs(ref_xyr)
    nbad = int(np.sum(np.any(err > tol, axis=1)))
    # vertex sums: oracle per-point derivatives weighted by gz, summed per triangle
    ref_tri = np.zeros((len(tris), 9))
    np.add.at(ref_tri, res["tri"][has], g[has, 3:] * gz[has, None])
    terr = np.abs(gtri - ref_tri)
    ttol = 10 * atol + rtol * np.abs(ref_tri) + rtol * 1e-2 * np.abs(ref_tri).max()
    nbad_t = int(np.sum(terr > ttol))
    ref_r = float(np.sum(g[has, 2] * gz[has])) if spec["kind"] != "composite" else float("nan")
    print(f"[grad-gpu] {label} {spec['kind']} fwd={fwd_prec} bwd={bwd_prec}: n={len(cl)} contacts={int(has.sum())} "
          f"max|d(xy,r)|={err.max():.2e} bad={nbad} max|dV|={terr.max():.2e} badV={nbad_t} "
          f"gr={gr:.6e} ref={ref_r:.6e}", flush=True)
    assert nbad == 0 and nbad_t == 0
    if spec["kind"] != "composite":
        assert abs(gr - ref_r) <= 1e-6 * (1 + abs(ref_r)) + (0 if bwd_prec == "fp64" else 1e-3 * np.sum(np.abs(g[has, 2])))


@pytest.mark.gpu
@pytest.mark.parametrize("bwd", ["mixed", "fp64"])
@pytest.mark.parametrize("name", list(CASES))
def test_device_backward_matches_oracle_gradient(gpu, grad_scene, name, bwd):
    tris, cl = grad_scene
    _check_device_vs_oracle(tris, CASES[name], cl, "exact64", bwd, "small")


@pytest.mark.gpu
def test_device_backward_config4_full(gpu):
    """The 524,288-triangle heightfield, BallCutter d=10, 1M uniform-random CL
    points: FP32 forward (mixed) and FP32 backward with FP64 edges against
    the FP64 analytic gradient at every point; upstream gradient N(0,1)."""
    tris = W.config1_mesh()
    cl = W.cl_random(1_000_000, 200.0, z0=-10.0, seed=44)
    _check_device_vs_oracle(tris, BALL10, cl, "mixed", "mixed", "C4")


@pytest.mark.gpu
def test_backward_vertex_sums_are_deterministic_over_long_segments(gpu):
    """Two large triangles under 200,000 random points: each triangle's terms
    span hundreds of 256-point blocks of the sorted order, so its sum is
    carried across blocks.  The sums match the oracle's per-point gradients
    summed per triangle, and repeated calls give the same bits (the sums
    have one fixed order, no atomics)."""
    import torch
    from opencamlib_amd import capi
    tris = np.array([[0, 0, 0, 100, 0, 10, 100, 100, 20], [0, 0, 0, 100, 100, 20, 0, 100, 10],
                     [200, 0, 0, 210, 0, 0, 200, 10, 0]], dtype=float)
    cl = W.cl_random(200_000, 100.0, z0=-10.0, seed=3)
    _check_device_vs_oracle(tris, BALL10, cl, "exact64", "fp64", "long-segments")
    gz = np.random.default_rng(5).normal(size=len(cl))
    dev = capi.DeviceMesh(tris)
    fw = _device_forward(dev, BALL10, cl, "mixed")
    runs = [_device_backward(dev, BALL10, fw, len(tris), gz, "mixed") for _ in range(3)]
    dev.close()
    for dz, gtri, gr in runs[1:]:
        assert np.array_equal(dz, runs[0][0])
        assert np.array_equal(gtri, runs[0][1])
        assert gr == runs[0][2]
    assert np.all(runs[0][1][2] == 0.0)  # the far triangle is under no point
    del torch


@pytest.mark.gpu
def test_autograd_function_routes_gradients(gpu, grad_scene):
    import torch
    from opencamlib_amd import capi
    from opencamlib_amd.autograd import drop_z
    tris, cl = grad_scene
    spec = BALL10
    dev = capi.DeviceMesh(tris)
    xy = torch.tensor(cl[:, :2], dtype=torch.float64, device="cuda", requires_grad=True)
    verts = torch.tensor(tris, dtype=torch.float64, device="cuda", requires_grad=True)
    rad = torch.tensor(spec["d"] / 2, dtype=torch.float64, device="cuda", requires_grad=True)
    z = drop_z(xy, verts, rad, dev, spec, z0=-10.0, precision="fp64")
    w = torch.linspace(-1, 1, len(cl), dtype=torch.float64, device="cuda")
    (z * w).sum().backward()
    fw = _device_forward(dev, spec, cl, "fp64")
    dz, gtri, gr = _device_backward(dev, spec, fw, len(tris), w.cpu().numpy(), "fp64")
    dev.close()
    assert torch.allclose(xy.grad.cpu(), torch.tensor(dz[:, :2] * w.cpu().numpy()[:, None]), atol=1e-12)
    assert np.allclose(verts.grad.cpu().numpy(), gtri, atol=1e-9)
    assert abs(rad.grad.item() - gr) < 1e-9 * (1 + abs(gr))
    # radius as a plain float, vertices without grad: only the CL gradient
    xy2 = xy.detach().clone().requires_grad_(True)
    dev = capi.DeviceMesh(tris)
    z2 = drop_z(xy2, verts.detach(), spec["d"] / 2, dev, spec, z0=-10.0, precision="fp64")
    (z2 * w).sum().backward()
    assert torch.equal(xy2.grad, xy.grad)
    # a radius that is not the spec's is refused
    with pytest.raises(ValueError, match="radius"):
        drop_z(xy2, verts.detach(), 1.0 + spec["d"] / 2, dev, spec, z0=-10.0, precision="fp64")
    dev.close()
