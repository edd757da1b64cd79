"""Batched localization (pt_batch_localization): stacks of uneven frame sizes, frame counts, densities
and fit classes (7 x 7 fit of sigma, 9 x 9 astigmatic fit of sigma_x / sigma_y) in one call. Each
stack's rows equal Localization of that stack alone with the same fast-fit schedule, and the FP64
oracle within the fit tolerance."""
import numpy as np
import pytest

from helpers import fit_errors, fit_mismatch_rows

pytestmark = pytest.mark.gpu
P2D = (7.0, 1.3, 1.3, 0.0)
P3D = (9.0, 1.4, 1.4, 0.0)
ASTIG = dict(astig=1, s0=1.2, c=250.0, d=400.0, zmin=-600.0, zmax=600.0)

# (side, frames, emitters per frame, fit class: 2 = 7x7 mode 2, 3 = 9x9 mode 3)
SPECS = [(128, 9, 8, 2), (256, 6, 60, 3), (512, 4, 200, 2), (128, 15, 30, 3), (1024, 2, 900, 2), (256, 11, 10, 2),
         (512, 3, 400, 3), (128, 1, 4, 3), (256, 5, 0, 2)]


@pytest.fixture(scope="module")
def stacks(pt_gpu):
    import torch
    out = []
    for k, (side, frames, em, cls) in enumerate(SPECS):
        d = torch.empty((frames, side, side), dtype=torch.uint16, device="cuda")
        cfg = ASTIG if cls == 3 else {}
        pt_gpu.synth_device(d, seed=0x5EED4000 + k, emitters=em, **cfg)
        thr = pt_gpu.autothreshold_device(d[0], ((9.0,) if cls == 3 else (7.0,)))
        mpp = max(64, 4 * em)
        out.append(dict(d_stack=d, threshold=thr, fit=cls, fit_params=(P3D if cls == 3 else P2D), max_per_plane=mpp))
    torch.cuda.synchronize()
    return out


def _fresh_outputs(stacks):
    import torch
    res = []
    for st in stacks:
        z = st["d_stack"].shape[0]
        st = dict(st)
        st["d_out"] = torch.full((z * st["max_per_plane"] * 18,), -3.0, dtype=torch.float64, device="cuda")
        res.append(st)
    return res


@pytest.fixture(autouse=True)
def options(pt_gpu):
    pt_gpu.set_option("fit_fp64", 0)
    yield
    pt_gpu.set_option("fit_fp64", 1)
    pt_gpu.set_option("fit_handover", 16)
    pt_gpu.set_option("batch_graph", 1)
    pt_gpu.set_option("batch_autotune", 1)


def test_batch_equals_each_stack_alone(pt_gpu, stacks):
    batch = _fresh_outputs(stacks)
    rows = pt_gpu.batch_localization(batch)
    sched = {(c["mode"], c["roi"]): c for c in pt_gpu.batch_schedules()}
    assert set(sched) == {(2, 7), (3, 9)}
    assert all(c["tuned"] for c in sched.values())
    for st, n in zip(batch, rows):
        c = sched[(st["fit"], int(st["fit_params"][0]))]
        pt_gpu.set_option("fit_handover", c["handover"])
        import torch
        alone = torch.zeros_like(st["d_out"])
        m = pt_gpu.localization_device(st["d_stack"], alone, st["threshold"], 0.0, st["fit"], st["fit_params"],
                                       st["max_per_plane"])
        assert n == m
        assert torch.equal(st["d_out"][: n * 18], alone[: m * 18])
        assert bool((st["d_out"][n * 18:] == -3.0).all())
    assert sum(rows) > 3000


def test_batch_matches_oracle(pt_gpu, oracle_mod, stacks):
    batch = _fresh_outputs(stacks)
    rows = pt_gpu.batch_localization(batch)
    for k in (1, 2, 6):
        st, n = batch[k], rows[k]
        got = st["d_out"][: n * 18].view(-1, 18).cpu().numpy()
        exp = oracle_mod.localization(st["d_stack"].cpu().numpy(), st["threshold"], 0.0, st["fit"], st["fit_params"],
                                      output_size=st["d_out"].numel())
        assert got.shape == exp.shape
        bad = fit_mismatch_rows(got, exp)
        assert bad.size == 0, (k, bad.size, fit_errors(got, exp))


def test_batch_graph_replays_equal_eager(pt_gpu, stacks):
    """The captured fit stage, replayed several times with eager work and synchronizations between
    the replays, gives the eager result every time."""
    import torch
    pt_gpu.set_option("batch_graph", 0)
    eager = _fresh_outputs(stacks)
    pt_gpu.batch_localization(eager)
    pt_gpu.set_option("batch_graph", 1)
    for rep in range(4):
        b = _fresh_outputs(stacks)
        pt_gpu.batch_localization(b)
        # Eager work between replays: a Localization of another stack and a device sync.
        other = torch.empty(stacks[2]["d_stack"].shape[0] * 800 * 18, dtype=torch.float64, device="cuda")
        pt_gpu.localization_device(stacks[2]["d_stack"], other, stacks[2]["threshold"], 0.0, 2, P2D, 800)
        torch.cuda.synchronize()
        for x, y in zip(b, eager):
            assert torch.equal(x["d_out"], y["d_out"]), rep


def test_batch_with_detection_only_stack(pt_gpu, stacks):
    """A stack with fit mode 0 in the batch gets the region rows of Localization with fit 0."""
    import torch
    mixed = _fresh_outputs(stacks[:3])
    mixed[1]["fit"] = 0
    rows = pt_gpu.batch_localization(mixed)
    alone = torch.zeros_like(mixed[1]["d_out"])
    m = pt_gpu.localization_device(mixed[1]["d_stack"], alone, mixed[1]["threshold"], 0.0, 0, P3D,
                                   mixed[1]["max_per_plane"])
    assert rows[1] == m and torch.equal(mixed[1]["d_out"][: m * 18], alone[: m * 18])
