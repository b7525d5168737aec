"""Parity of the HIP batch drop-cutter with the FP64 oracle (GPU).

Bar: exact64 reproduces the oracle's z bit for bit (Bull: the device atan/sin
of the offset-ellipse solve differ from libm by an ulp, so 1e-12) and the
oracle's contact triangle; fp64 and mixed reproduce every contact type, with
|dz| <= 1e-9 for fp64 and <= 1e-6 mm for mixed (the winner is re-evaluated by
the exact FP64 code); fp32 is
reported against a 1e-3 mm bound and a type-agreement floor.  Where the
device reports a different triangle than the oracle, that triangle must
realise the same contact height (a shared edge or vertex seen from the
neighbouring triangle).
"""
import math

import numpy as np
import pytest

from cutters import ALL, BALL10, BULL10, ball_cone, cyl_cone
from opencamlib_amd import capi, workloads as W
from oracle import oracle as O

pytestmark = pytest.mark.gpu

MODES = ["exact64", "fp64", "mixed", "fp32"]


def oracle_drop(tris, spec, cl, bucket=1):
    return O.Mesh(tris).batch_drop(O.cutter_from_spec(spec), cl, search_mode=1, bucket=bucket)


def check_alt_triangles(tris, spec, cl, got, ref, mask, ztol=1e-9):
    cut = O.cutter_from_spec(spec)
    bad = np.nonzero(mask & (got["tri"] != ref["tri"]))[0]
    for i in bad[:200]:
        alt = O.drop_triangle(cut, tris[got["tri"][i]], cl[i, 0], cl[i, 1], ref["z"][i] - 1e-7)
        assert alt["type"] == ref["type"][i], (i, alt, ref["type"][i])
        assert abs(alt["z"] - ref["z"][i]) < ztol
    return len(bad)


def compare(tris, spec, cl, got, ref, mode, label=""):
    dz = np.abs(got["z"] - ref["z"])
    ty_mis = int(np.sum(got["type"] != ref["type"]))
    p99 = float(np.percentile(dz, 99)) if len(dz) else 0.0
    print(f"[parity] {label} {spec['kind']} {mode}: n={len(cl)} max|dz|={dz.max() if len(dz) else 0:.3e} "
          f"p99|dz|={p99:.3e} type_mismatch={ty_mis} stats={got['stats']}", flush=True)
    if mode == "exact64":
        tol = 1e-12 if spec["kind"] == "bull" else 0.0
        assert dz.max() <= tol
        assert ty_mis == 0
        if spec["kind"] != "bull":
            assert np.array_equal(got["tri"], ref["tri"])
            assert np.array_equal(got["cc"], ref["cc"])
        else:
            check_alt_triangles(tris, spec, cl, got, ref, np.ones(len(cl), bool))
    elif mode in ("fp64", "mixed"):
        # mixed: two same-type contacts closer than the FP32 error band are
        # ordered in FP32 and the chosen one re-evaluated in FP64, so z can
        # sit below the true maximum by at most their gap (bound 1e-6 mm,
        # well inside the 1 um target); the type is always exact.
        ztol = 1e-9 if mode == "fp64" else 1e-6
        assert dz.max() <= ztol
        assert ty_mis == 0
        check_alt_triangles(tris, spec, cl, got, ref, np.ones(len(cl), bool), ztol)
        lifted = ref["type"] != 0
        cc_off = np.max(np.abs(got["cc"] - ref["cc"]), axis=1) > 1e-6
        n_cc = int(np.sum(cc_off & lifted))
        print(f"[parity]   cc points differing: {n_cc}", flush=True)
        if mode == "fp64":
            assert n_cc == 0
        else:
            # a different CC is an equal-type contact within ztol of the
            # maximum (z and type were checked above); it must stay rare
            assert n_cc <= max(2, 1e-4 * len(cl))
    else:
        # FP32 only, no refinement: reported, with loose floors.  Cone tips on
        # a CL grid aligned with the mesh grid sit exactly on triangle edges,
        # where FP32 rounding decides between equal-height contacts.
        assert dz.max() <= 1e-3
        assert ty_mis <= (0.1 if spec["kind"] == "cone" else 0.01) * len(cl)


@pytest.fixture(scope="module")
def c0():
    return W.config0_mesh(), W.config0_cl()


@pytest.mark.parametrize("mode", MODES)
@pytest.mark.parametrize("name", ["cyl", "ball", "bull", "cone", "cylball"])
def test_config0_all_cutters(gpu, c0, name, mode):
    tris, cl = c0
    spec = ALL[name]
    ref = oracle_drop(tris, spec, cl)
    mesh = capi.DeviceMesh(tris)
    got = mesh.drop(spec, cl, mode)
    mesh.close()
    compare(tris, spec, cl, got, ref, mode, "C0")


@pytest.mark.parametrize("mode", ["exact64", "mixed"])
@pytest.mark.parametrize("name", ["cylcone", "ballcone"])
def test_composites_with_cone_parts(gpu, name, mode):
    """Composites whose cone part sits below the first cutter (negative
    offsets: contacts can rise above a triangle's max z): the exact path's
    running-height skip must leave every result bit-equal to the oracle."""
    spec = cyl_cone(2.0, 10.0, math.pi / 4) if name == "cylcone" else ball_cone(4.0, 10.0, math.pi / 4)
    tris = W.heightfield(32, 50.0, bumps=3, steps=3, seed=23)
    cl = W.cl_grid(70, 70, 0.7, z0=-10.0, origin=(0.3, 0.4))
    ref = oracle_drop(tris, spec, cl)
    mesh = capi.DeviceMesh(tris)
    got = mesh.drop(spec, cl, mode)
    mesh.close()
    assert np.sum(ref["type"] != 0) > len(cl) // 2
    assert np.array_equal(got["z"], ref["z"])
    assert np.array_equal(got["type"], ref["type"])
    assert np.array_equal(got["tri"], ref["tri"])
    assert np.array_equal(got["cc"], ref["cc"])


@pytest.fixture(scope="module")
def c1_mesh():
    return W.config1_mesh()


@pytest.fixture(scope="module")
def c1_device(gpu, c1_mesh):
    m = capi.DeviceMesh(c1_mesh)
    yield m
    m.close()


@pytest.fixture(scope="module")
def c1_oracle(c1_mesh):
    return O.Mesh(c1_mesh)


@pytest.mark.parametrize("mode", MODES)
@pytest.mark.parametrize("spec", [BALL10, BULL10], ids=["ball", "bull"])
def test_config1_subset(c1_mesh, c1_device, c1_oracle, spec, mode):
    """524,288-triangle mesh with 20 steps (vertical walls); a seeded
    50,000-point subset of the 1000x1000 CL grid."""
    cl = W.subsample(W.config1_cl(), 50_000, seed=17)
    ref = c1_oracle.batch_drop(O.cutter_from_spec(spec), cl, search_mode=1)
    got = c1_device.drop(spec, cl, mode)
    compare(c1_mesh, spec, cl, got, ref, mode, "C1-50k")


@pytest.mark.parametrize("name", ["cyl", "cone", "cylball"])
def test_config1_other_cutters_mixed(c1_mesh, c1_device, c1_oracle, name):
    spec = ALL[name]
    cl = W.subsample(W.config1_cl(), 20_000, seed=18)
    ref = c1_oracle.batch_drop(O.cutter_from_spec(spec), cl, search_mode=1)
    got = c1_device.drop(spec, cl, "mixed")
    compare(c1_mesh, spec, cl, got, ref, "mixed", "C1-20k")


def test_config1_full_ball_mixed(c1_mesh, c1_device, c1_oracle):
    """All 1,000,000 CL points, BallCutter d=10: the bench workload."""
    cl = W.config1_cl()
    ref = c1_oracle.batch_drop(O.cutter_from_spec(BALL10), cl, search_mode=1)
    got = c1_device.drop(BALL10, cl, "mixed")
    compare(c1_mesh, BALL10, cl, got, ref, "mixed", "C1-full")


# ------------------------------------------------------------ edge cases
def test_empty_mesh_and_empty_batch(gpu):
    spec = BALL10
    m = capi.DeviceMesh(np.zeros((0, 9)))
    cl = W.cl_grid(5, 5, 1.0, z0=-3.0)
    for mode in MODES:
        got = m.drop(spec, cl, mode)
        assert np.all(got["z"] == -3.0) and np.all(got["type"] == 0)
        assert np.all(got["tri"] == -1)
    m.close()
    m = capi.DeviceMesh(W.heightfield(4, 10.0, bumps=1, seed=1))
    got = m.drop(spec, np.zeros((0, 3)), "mixed")
    assert got["z"].shape == (0,)
    m.close()


@pytest.mark.parametrize("name", ["cyl", "ball", "bull", "cone", "cylball"])
def test_degenerate_and_ragged_meshes(gpu, name):
    """Vertical triangles, zero-area slivers, repeated vertices, a triangle
    count that does not fill the last 16-triangle box, CL points far outside
    the mesh and CL points already above it."""
    rng = np.random.default_rng(5)
    tris = [W.heightfield(5, 12.0, bumps=1, seed=3)]
    vert = np.array([[3, 3, 0, 7, 3, 0, 5, 3, 9], [3, 3, 0, 3, 7, 0, 3, 5, 8]], dtype=float)
    sliver = np.array([[1, 1, 2, 5, 5, 6, 9, 9, 10], [4, 4, 4, 4, 4, 4, 6, 6, 1]], dtype=float)
    rand = rng.uniform(0, 12, (7, 9))
    tris = np.concatenate(tris + [vert, sliver, rand])
    assert len(tris) % 16 != 0
    cl = np.concatenate([W.cl_grid(60, 60, 0.25, z0=-20.0, origin=(-1.5, -1.5)),
                         np.array([[100.0, 100.0, -5.0], [-50.0, 4.0, -1.0], [5.0, 5.0, 50.0]])])
    spec = ALL[name]
    ref = oracle_drop(tris, spec, cl)
    m = capi.DeviceMesh(tris)
    for mode in ("exact64", "fp64", "mixed"):
        got = m.drop(spec, cl, mode)
        compare(tris, spec, cl, got, ref, mode, "degenerate")
    m.close()


@pytest.mark.parametrize("bucket", [1, 3, 16])
def test_bucket_size_changes_tie_order_like_reference(gpu, bucket):
    """Operation::setBucketSize changes the KD leaf order and with it the
    equal-height winner; the device follows the oracle for every bucket."""
    tris = W.heightfield(32, 40.0, bumps=0, noise=0.0, seed=0)  # flat: every contact ties
    tris[:, 2::3] += 1.0
    cl = W.cl_grid(50, 50, 0.8, z0=-5.0)
    spec = ALL["cyl"]
    ref = oracle_drop(tris, spec, cl, bucket=bucket)
    m = capi.DeviceMesh(tris, bucket=bucket)
    got = m.drop(spec, cl, "exact64")
    assert np.array_equal(got["z"], ref["z"])
    assert np.array_equal(got["tri"], ref["tri"])
    assert np.array_equal(got["type"], ref["type"])
    m.close()


@pytest.mark.parametrize("shift", [(1000.0, -2000.0, 50.0), (1e5, 1e5, -300.0)])
def test_translation_equivariance_far_from_origin(gpu, shift):
    """Coordinates far from the origin: the FP32 selection works in box-local
    frames, so a translated scene gives the translated heights and the same
    contacts."""
    tris = W.heightfield(40, 50.0, bumps=3, steps=4, seed=8)
    cl = W.cl_grid(120, 120, 0.4, z0=-10.0)
    t2 = tris.copy()
    t2[:, 0::3] += shift[0]
    t2[:, 1::3] += shift[1]
    t2[:, 2::3] += shift[2]
    cl2 = cl + np.array(shift)
    for spec in (BALL10, BULL10):
        ref = oracle_drop(t2, spec, cl2)
        m = capi.DeviceMesh(t2)
        got = m.drop(spec, cl2, "mixed")
        m.close()
        compare(t2, spec, cl2, got, ref, "mixed", f"shift{shift}")
        base = oracle_drop(tris, spec, cl)
        assert np.max(np.abs(ref["z"] - shift[2] - base["z"])) < 1e-8


def test_device_pointer_entry_matches_host_entry(gpu):
    import torch
    tris = W.heightfield(24, 30.0, bumps=2, steps=2, seed=12)
    cl = W.cl_random(5000, 30.0, seed=4)
    m = capi.DeviceMesh(tris)
    host = m.drop(BULL10, cl, "mixed")
    d_cl = torch.from_numpy(cl).cuda()
    z = torch.empty(len(cl), dtype=torch.float64, device="cuda")
    cc = torch.empty((len(cl), 3), dtype=torch.float64, device="cuda")
    ty = torch.empty(len(cl), dtype=torch.uint8, device="cuda")
    tri = torch.empty(len(cl), dtype=torch.int32, device="cuda")
    ft = torch.empty(len(cl), dtype=torch.uint8, device="cuda")
    stream = torch.cuda.current_stream().cuda_stream
    m.drop_device(BULL10, d_cl.data_ptr(), len(cl), "mixed", z.data_ptr(), cc.data_ptr(), ty.data_ptr(),
                  tri.data_ptr(), ft.data_ptr(), stream=stream)
    torch.cuda.synchronize()
    assert np.array_equal(z.cpu().numpy(), host["z"])
    assert np.array_equal(ty.cpu().numpy(), host["type"])
    assert np.array_equal(tri.cpu().numpy().astype(np.int64), host["tri"])
    assert np.array_equal(cc.cpu().numpy(), host["cc"])
    m.close()
