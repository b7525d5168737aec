"""Batched drop-cutter jobs (ocl_hip_batch_*): one persistent selection
launch over a device work queue, optionally captured in a hipGraph.

Bar: every job's outputs equal, bit for bit, what ocl_hip_drop_device gives
for that job alone (same Hilbert resolution); the per-job-launch baseline
and every graph replay give the same bits; other Hilbert resolutions agree
with the FP64 oracle to the single-job bar (types exact, |dz| <= 1e-6 mm).
"""
import numpy as np
import pytest

from opencamlib_amd import capi, workloads as W
from oracle import oracle as O

pytestmark = pytest.mark.gpu
FIELDS = ("z", "cc", "type", "tri", "feat")


def small_jobs(seed=9, njobs=9):
    return W.config3_jobs(njobs, seed=seed, tri_range=(200, 20_000), cl_range=(100, 6_000))


class DeviceJobs:
    """Device copies of the jobs' CL arrays and output buffers."""

    def __init__(self, jobs):
        import torch
        self.jobs = jobs
        self.meshes = [capi.DeviceMesh(j["tris"]) for j in jobs]
        self.cl = [torch.tensor(j["cl"], dtype=torch.float64, device="cuda") for j in jobs]
        self.out = [self._alloc(len(j["cl"])) for j in jobs]

    @staticmethod
    def _alloc(n):
        import torch
        return {"z": torch.full((n,), float("nan"), dtype=torch.float64, device="cuda"),
                "cc": torch.full((n, 3), float("nan"), dtype=torch.float64, device="cuda"),
                "type": torch.full((n,), 255, dtype=torch.uint8, device="cuda"),
                "tri": torch.full((n,), -7, dtype=torch.int32, device="cuda"),
                "feat": torch.full((n,), 255, dtype=torch.uint8, device="cuda")}

    def ctypes_jobs(self):
        items = []
        for m, j, c, o in zip(self.meshes, self.jobs, self.cl, self.out):
            items.append({"mesh": m, "spec": j["spec"], "cl": c.data_ptr(), "n": len(j["cl"]),
                          **{k: o[k].data_ptr() for k in FIELDS}})
        return capi.make_jobs(items)

    def reset_outputs(self):
        for o in self.out:
            o["z"].fill_(float("nan"))
            o["cc"].fill_(float("nan"))
            o["type"].fill_(255)
            o["tri"].fill_(-7)
            o["feat"].fill_(255)

    def snapshot(self):
        import torch
        torch.cuda.synchronize()
        return [{k: o[k].cpu().numpy().copy() for k in FIELDS} for o in self.out]

    def single(self, precision):
        """Each job alone through ocl_hip_drop_device."""
        import torch
        res = []
        for m, j, c in zip(self.meshes, self.jobs, self.cl):
            o = self._alloc(len(j["cl"]))
            m.drop_device(j["spec"], c.data_ptr(), len(j["cl"]), precision, *(o[k].data_ptr() for k in FIELDS))
            torch.cuda.synchronize()
            res.append({k: o[k].cpu().numpy() for k in FIELDS})
        return res

    def close(self):
        for m in self.meshes:
            m.close()


def same_bits(a, b, fields=FIELDS):
    for x, y in zip(a, b):
        for k in fields:
            if not np.array_equal(x[k], y[k], equal_nan=(k in ("z", "cc"))):
                return False
    return True


@pytest.fixture(scope="module")
def djobs(gpu):
    d = DeviceJobs(small_jobs())
    yield d
    d.close()


@pytest.mark.parametrize("precision", ["mixed", "fp64", "fp32", "exact64"])
def test_batch_equals_single_jobs(djobs, precision):
    ref = djobs.single(precision)
    kinds = sorted({j["spec"]["kind"] for j in djobs.jobs})
    for tune in ({}, {"per_job_launch": 1}):
        djobs.reset_outputs()
        b = capi.Batch(djobs.ctypes_jobs(), precision, tune)
        st = b.run()
        per_job = [b.job_stats(j) for j in range(len(djobs.jobs))]
        got = djobs.snapshot()
        b.close()
        # the per-job counters add up to the batch's
        assert [p["points"] for p in per_job] == [len(j["cl"]) for j in djobs.jobs]
        assert sum(p["pairs"] for p in per_job) == st["pairs"]
        assert sum(p["escalated"] for p in per_job) == st["escalated"]
        print(f"[batch] {precision} tune={tune} jobs={len(got)} kinds={kinds} stats={st}", flush=True)
        fields = FIELDS if precision != "fp32" else ("z", "type", "tri", "feat")
        assert same_bits(got, ref, fields)


def test_graph_replays_match_eager_and_follow_inputs(djobs):
    import torch
    b = capi.Batch(djobs.ctypes_jobs(), "mixed", {"use_graph": 1})
    djobs.reset_outputs()
    b.run()
    first = djobs.snapshot()
    assert same_bits(first, djobs.single("mixed"))
    for r in range(3):
        djobs.reset_outputs()
        _ = torch.ones(1000, device="cuda").sum().item()  # eager work between replays
        b.run()
        assert same_bits(djobs.snapshot(), first), r
    # the replay re-reads the CL arrays: move every job's points
    for c in djobs.cl:
        c[:, 0] += 0.37
    djobs.reset_outputs()
    b.run()
    moved = djobs.snapshot()
    ref = djobs.single("mixed")
    for c in djobs.cl:
        c[:, 0] -= 0.37
    b.close()
    assert same_bits(moved, ref)


@pytest.mark.parametrize("hb", [6, 8, 10])
def test_other_hilbert_resolutions_match_oracle(djobs, hb):
    djobs.reset_outputs()
    b = capi.Batch(djobs.ctypes_jobs(), "mixed", {"hilbert_bits": hb, "blocks_per_cu": 2, "use_graph": 1})
    b.run()
    got = djobs.snapshot()
    b.close()
    for j, g in zip(djobs.jobs, got):
        ref = O.Mesh(j["tris"]).batch_drop(O.cutter_from_spec(j["spec"]), j["cl"], search_mode=1)
        assert np.array_equal(g["type"], ref["type"].astype(np.uint8))
        assert np.max(np.abs(g["z"] - ref["z"])) <= 1e-6


def test_autotune_returns_a_working_tune(djobs):
    best, ms = capi.autotune(djobs.ctypes_jobs(), "mixed", trials=1)
    print(f"[autotune] best={best} ms={ms:.3f}", flush=True)
    assert best["hilbert_bits"] in (7, 8, 9) and best["blocks_per_cu"] in (2, 4, 8) and ms > 0
    djobs.reset_outputs()
    b = capi.Batch(djobs.ctypes_jobs(), "mixed", best)
    b.run()
    got = djobs.snapshot()
    b.close()
    assert all(np.all(np.isfinite(g["z"])) for g in got)


def test_batch_errors(djobs):
    jobs = djobs.ctypes_jobs()
    with pytest.raises(capi.OclError, match="hilbert_bits"):
        capi.Batch(jobs, "mixed", {"hilbert_bits": 12})
    with pytest.raises(capi.OclError, match="at least one job"):
        capi.Batch((capi.DropJob * 0)(), "mixed")


def test_batches_keep_their_state_when_the_mesh_is_reused(djobs):
    """A batch keeps its own selection records and workspace.  Two batches
    on the same meshes with different cutters, a second mesh appearing twice
    in one batch with two cutters, and single calls with a larger point set
    in between leave every replay bit-equal to the single calls; a mesh used
    by a live batch cannot be destroyed."""
    import torch
    swap = {"cyl": "ball", "ball": "bull", "bull": "cone", "cone": "cyl", "composite": "ball"}
    from cutters import ALL
    other_specs = [ALL[swap[j["spec"]["kind"]]] for j in djobs.jobs]
    b1 = capi.Batch(djobs.ctypes_jobs(), "mixed", {"use_graph": 1})
    ref1 = djobs.single("mixed")
    # batch 2: the same meshes with other cutters (rebuilds each mesh's records)
    alt = DeviceJobs.__new__(DeviceJobs)
    alt.jobs = [dict(j, spec=s) for j, s in zip(djobs.jobs, other_specs)]
    alt.meshes, alt.cl = djobs.meshes, djobs.cl
    alt.out = [DeviceJobs._alloc(len(j["cl"])) for j in alt.jobs]
    b2 = capi.Batch(alt.ctypes_jobs(), "mixed", {"use_graph": 1})
    ref2 = alt.single("mixed")
    # a single call with 4x the points of every job (grows each mesh's workspace)
    for m, j, c in zip(djobs.meshes, djobs.jobs, djobs.cl):
        big = torch.cat([c] * 4)
        o = DeviceJobs._alloc(len(big))
        m.drop_device(j["spec"], big.data_ptr(), len(big), "mixed", *(o[k].data_ptr() for k in FIELDS))
    torch.cuda.synchronize()
    for r in range(2):
        djobs.reset_outputs()
        b1.run()
        assert same_bits(djobs.snapshot(), ref1), ("batch 1", r)
        alt.reset_outputs()
        b2.run()
        assert same_bits(alt.snapshot(), ref2), ("batch 2", r)
    # one mesh twice in a batch, with two cutters
    two = DeviceJobs.__new__(DeviceJobs)
    two.jobs = [djobs.jobs[0], dict(djobs.jobs[0], spec=other_specs[0])]
    two.meshes = [djobs.meshes[0], djobs.meshes[0]]
    two.cl = [djobs.cl[0], djobs.cl[0]]
    two.out = [DeviceJobs._alloc(len(j["cl"])) for j in two.jobs]
    b3 = capi.Batch(two.ctypes_jobs(), "mixed", {"use_graph": 1})
    b3.run()
    assert same_bits(two.snapshot(), two.single("mixed"))
    with pytest.raises(capi.OclError, match="live batch"):
        djobs.meshes[0].close()
    for b in (b1, b2, b3):
        b.close()
