"""The allocation-free drop entry point (ocl_hip_drop_device_noalloc)
against ocl_hip_drop_device, through the C ABI (GPU).

Bar: bit-equal outputs (z, CC point, CC type, triangle, feature; FP32 mode
has no CC point) for every cutter kind and precision, on two calls in a
row, the second with a smaller input; point sets above the limit and meshes
without the cutter's selection records are refused before any device work.
"""
import numpy as np
import pytest

from cutters import ALL
from opencamlib_amd import capi, workloads as W

pytestmark = pytest.mark.gpu
FIELDS = ("z", "cc", "type", "tri", "feat")
LIMIT = 1 << 19


def alloc(n):
    import torch
    return {"z": torch.full((n,), float("nan"), dtype=torch.float64, device="cuda"),
            "cc": torch.full((n, 3), float("nan"), dtype=torch.float64, device="cuda"),
            "type": torch.full((n,), 255, dtype=torch.uint8, device="cuda"),
            "tri": torch.full((n,), -7, dtype=torch.int32, device="cuda"),
            "feat": torch.full((n,), 255, dtype=torch.uint8, device="cuda")}


def run(mesh, spec, cl, precision, noalloc):
    import torch
    o = alloc(cl.shape[0])
    args = (spec, cl.data_ptr(), cl.shape[0], precision, *(o[k].data_ptr() for k in FIELDS))
    if noalloc:
        mesh.drop_device_noalloc(*args)
    else:
        mesh.drop_device(*args)
    torch.cuda.synchronize()
    return {k: o[k].cpu().numpy() for k in FIELDS}


def same(a, b):
    return all(np.array_equal(a[k], b[k], equal_nan=k in ("z", "cc")) for k in FIELDS)


@pytest.fixture(scope="module")
def scene(gpu):
    import torch
    tris = W.heightfield(64, 100.0, bumps=3, steps=2, seed=8)
    big = W.cl_grid(450, 450, 0.22, z0=-10.0)  # 202,500 points
    small = big[:53_000].copy()
    small[:, :2] += 0.11
    return tris, torch.tensor(big, device="cuda"), torch.tensor(small, device="cuda")


@pytest.mark.parametrize("precision", ["mixed", "fp64", "fp32", "exact64"])
@pytest.mark.parametrize("name", ["ball", "bull", "cone", "cylball"])
def test_noalloc_equals_drop_device_on_two_calls(scene, name, precision):
    tris, big, small = scene
    spec = ALL[name]
    mesh = capi.DeviceMesh(tris)
    for label, cl in (("first", big), ("second, smaller", small)):
        ref = run(mesh, spec, cl, precision, noalloc=False)  # builds the cutter's records
        got = run(mesh, spec, cl, precision, noalloc=True)
        print(f"[noalloc] {name} {precision} {label} call: n={cl.shape[0]} "
              f"max|dz|={np.nanmax(np.abs(got['z'] - ref['z'])):.3e}", flush=True)
        assert same(got, ref), (name, precision, label)
    mesh.close()


def test_noalloc_refuses_what_it_cannot_do(scene):
    import torch
    tris, big, small = scene
    mesh = capi.DeviceMesh(tris)
    out = alloc(8)
    ptrs = [out[k].data_ptr() for k in FIELDS]
    # more points than the scratch holds: refused before any device work
    with pytest.raises(capi.OclError, match="at most 524288"):
        mesh.drop_device_noalloc(ALL["ball"], big.data_ptr(), LIMIT + 1, "mixed", *ptrs)
    # no selection records for this cutter on this mesh yet
    with pytest.raises(capi.OclError, match="selection records"):
        mesh.drop_device_noalloc(ALL["bull"], small.data_ptr(), 8, "fp64", *ptrs)
    # exact64 and composites need no records: works on a fresh mesh
    got = run(mesh, ALL["ball"], small, "exact64", noalloc=True)
    ref = run(mesh, ALL["ball"], small, "exact64", noalloc=False)
    assert same(got, ref)
    # exactly the limit is accepted
    n = LIMIT
    cl = torch.cat([big, big, big])[:n].contiguous()
    run(mesh, ALL["ball"], cl[:1000].contiguous(), "mixed", noalloc=False)
    got = run(mesh, ALL["ball"], cl, "mixed", noalloc=True)
    ref = run(mesh, ALL["ball"], cl, "mixed", noalloc=False)
    assert same(got, ref)
    mesh.close()
