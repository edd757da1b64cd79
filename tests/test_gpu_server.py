"""Serving loop (pt_server_*): requests from several simulated cameras, admitted onto engine slots
under an HBM budget, give the same rows as the same Localization run alone."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu
FIT_PARAMS = (7.0, 1.3, 1.3, 0.0)


@pytest.fixture
def cameras(pt_gpu):
    """Six chunks: 256^2 and 512^2 frames, 3 to 40 frames, different densities and seeds."""
    import torch
    specs = [(256, 12, 60, 1), (512, 5, 200, 2), (256, 40, 30, 3), (512, 3, 400, 4), (256, 7, 120, 5), (512, 9, 90, 6)]
    cams = []
    for size, frames, emitters, seed in specs:
        d = torch.empty((frames, size, size), dtype=torch.uint16, device="cuda")
        pt_gpu.synth_device(d, seed=0x5E0000 + seed, emitters=emitters)
        thr = pt_gpu.autothreshold_device(d[0])
        cams.append((d, thr, 4 * emitters + 64))
    torch.cuda.synchronize()
    return cams


def _alone(pt_gpu, cams, ratio, fit):
    import torch
    out = []
    for d, thr, mpp in cams:
        o = torch.zeros(d.shape[0] * mpp * 18, dtype=torch.float64, device="cuda")
        n = pt_gpu.localization_device(d, o, thr, ratio, fit, FIT_PARAMS, mpp)
        out.append(o[: n * 18].view(-1, 18).cpu().numpy())
    return out


@pytest.mark.parametrize("fp64,ratio,fit", [(1, 0.0, 2), (0, 0.0, 2), (0, 10.0, 4)])
def test_server_results_equal_alone(pt_gpu, cameras, fp64, ratio, fit):
    import torch
    pt_gpu.set_option("fit_fp64", fp64)
    try:
        exp = _alone(pt_gpu, cameras, ratio, fit)
        with pt_gpu.Server(slots=3, budget_bytes=8 << 30, chunk_pixels=1 << 22) as srv:
            outs, ids = [], []
            for rep in range(2):                       # every chunk twice, all queued at once
                for d, thr, mpp in cameras:
                    o = torch.full((d.shape[0] * mpp * 18,), -7.0, dtype=torch.float64, device="cuda")
                    outs.append(o)
                    ids.append(srv.submit(d, o, thr, ratio, fit, FIT_PARAMS, mpp))
            slots = set()
            for k, rid in enumerate(ids):
                n, t = srv.wait(rid)
                slots.add(t[3])
                assert t[0] <= t[1] <= t[2]
                got = outs[k][: n * 18].view(-1, 18).cpu().numpy()
                e = exp[k % len(cameras)]
                assert got.shape == e.shape and got.shape[0] > 0
                assert np.array_equal(got, e, equal_nan=True), f"request {k}"
            assert len(slots) > 1                      # the requests were spread over the slots
    finally:
        pt_gpu.set_option("fit_fp64", 1)


def test_server_budget_admits_one_at_a_time(pt_gpu, cameras):
    """With a budget below two requests' HBM (frames, rows and engine work space), no two requests are
    in flight together; a request above the budget is refused."""
    import torch
    d, thr, mpp = cameras[1]
    cap = d.shape[0] * mpp * 18
    with pt_gpu.Server(slots=1, budget_bytes=1 << 40, chunk_pixels=1 << 22) as probe:
        req_bytes = probe.request_bytes(d.shape[0], d.shape[1], d.shape[2], cap, mpp)
    assert req_bytes > d.numel() * 2 + cap * 8
    with pt_gpu.Server(slots=3, budget_bytes=req_bytes + req_bytes // 2, chunk_pixels=1 << 22) as srv:
        assert srv.request_bytes(d.shape[0], d.shape[1], d.shape[2], cap, mpp) == req_bytes
        outs = [torch.zeros(cap, dtype=torch.float64, device="cuda") for _ in range(4)]
        ids = [srv.submit(d, o, thr, 0.0, 2, FIT_PARAMS, mpp) for o in outs]
        spans = [srv.wait(r)[1] for r in ids]
        spans.sort(key=lambda t: t[1])
        for a, b in zip(spans, spans[1:]):
            assert b[1] >= a[2], (a, b)
        # A request whose output alone exceeds the budget is refused at submission (its buffer is
        # never touched, so the C entry is given a capacity larger than the buffer behind it).
        params = np.asarray(FIT_PARAMS, np.float64)
        over = (req_bytes + req_bytes // 2) // 8 + 1
        assert pt_gpu.lib().pt_server_submit(srv._h, d.data_ptr(), d.shape[0], d.shape[1], d.shape[2], thr, 0.0, 2,
                                             params.ctypes.data, 4, outs[0].data_ptr(), over, mpp) == -1
        first = outs[0].cpu().numpy()
        for o in outs[1:]:
            assert np.array_equal(o.cpu().numpy(), first)


def test_server_budget_at_the_limit(pt_gpu, cameras):
    """A budget of exactly one request's HBM runs requests of two sizes one at a time (idle slots give
    their work space back when the other size needs the room); one byte less refuses the request."""
    import torch
    (d1, thr1, mpp1), (d2, thr2, mpp2) = cameras[1], cameras[2]
    caps = [d1.shape[0] * mpp1 * 18, d2.shape[0] * mpp2 * 18]
    with pt_gpu.Server(slots=1, budget_bytes=1 << 40, chunk_pixels=1 << 22) as probe:
        rb = [probe.request_bytes(d.shape[0], d.shape[1], d.shape[2], c, m)
              for (d, _, m), c in zip([cameras[1], cameras[2]], caps)]
    budget = max(rb)
    exp = [pt_gpu.localization(d.cpu().numpy(), t, 0.0, 2, FIT_PARAMS, output_size=c)
           for (d, t, _), c in zip([cameras[1], cameras[2]], caps)]
    with pt_gpu.Server(slots=2, budget_bytes=budget, chunk_pixels=1 << 22) as srv:
        jobs = []
        for k in range(6):
            i = k % 2
            d, thr, mpp = cameras[1 + i]
            o = torch.zeros(caps[i], dtype=torch.float64, device="cuda")
            jobs.append((i, o, srv.submit(d, o, thr, 0.0, 2, FIT_PARAMS, mpp)))
        spans = []
        for i, o, rid in jobs:
            n, t = srv.wait(rid)
            spans.append(t)
            assert np.array_equal(o[: n * 18].view(-1, 18).cpu().numpy(), exp[i])
        spans.sort(key=lambda t: t[1])
        for a, b in zip(spans, spans[1:]):
            assert b[1] >= a[2], (a, b)
    with pt_gpu.Server(slots=1, budget_bytes=rb[0] - 1, chunk_pixels=1 << 22) as srv:
        d, thr, mpp = cameras[1]
        o = torch.zeros(caps[0], dtype=torch.float64, device="cuda")
        with pytest.raises(RuntimeError):
            srv.submit(d, o, thr, 0.0, 2, FIT_PARAMS, mpp)


def test_server_work_space_estimate_bounds_allocation(pt_gpu, cameras):
    """The device memory a slot allocates for a request stays within the admission's estimate."""
    import torch
    d, thr, mpp = cameras[1]
    cap = d.shape[0] * mpp * 18
    o = torch.zeros(cap, dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    with pt_gpu.Server(slots=1, budget_bytes=1 << 40, chunk_pixels=1 << 22) as srv:
        ws = srv.request_bytes(d.shape[0], d.shape[1], d.shape[2], cap, mpp) - d.numel() * 2 - cap * 8
        free0 = torch.cuda.mem_get_info()[0]
        srv.wait(srv.submit(d, o, thr, 0.0, 2, FIT_PARAMS, mpp))
        used = free0 - torch.cuda.mem_get_info()[0]
    assert used <= ws, (used, ws)


def test_served_dense_cameras_match_oracle(pt_gpu, oracle_mod):
    """Requests at the served load's density, in flight together on two slots with the fast fit and
    the 20-iteration cap: a 2048^2 camera chunk at 5000 emitters per frame (every frame flagged and
    flooded whole) and a 256^2 one at 1500; every served row within the fit tolerance of the oracle."""
    import torch
    from helpers import fit_errors, fit_mismatch_rows
    specs = [(2048, 3, 5000, 0x5EED6001), (256, 10, 1500, 0x5EED6002)]
    chunks = []
    for side, frames, emitters, seed in specs:
        d = torch.empty((frames, side, side), dtype=torch.uint16, device="cuda")
        pt_gpu.synth_device(d, seed=seed, emitters=emitters)
        st = d.cpu().numpy()
        chunks.append((d, st, oracle_mod.autothreshold(st[0]), int(emitters * 1.25) + 64))
    pt_gpu.set_option("fit_fp64", 0)
    pt_gpu.set_option("max_iter", 20)
    try:
        with pt_gpu.Server(slots=2, budget_bytes=64 << 30, chunk_pixels=1 << 25) as srv:
            outs, ids = [], []
            for d, st, thr, mpp in chunks:
                o = torch.full((d.shape[0] * mpp * 18,), -7.0, dtype=torch.float64, device="cuda")
                outs.append(o)
                ids.append(srv.submit(d, o, thr, 0.0, 2, FIT_PARAMS, mpp))
            got = []
            for rid, o in zip(ids, outs):
                n, _ = srv.wait(rid)
                got.append(o[: n * 18].view(-1, 18).cpu().numpy())
    finally:
        pt_gpu.set_option("fit_fp64", 1)
        pt_gpu.set_option("max_iter", 75)
    for (d, st, thr, mpp), g in zip(chunks, got):
        exp = oracle_mod.localization(st, thr, 0.0, 2, FIT_PARAMS, output_size=st.shape[0] * mpp * 18, max_iter=20)
        assert g.shape == exp.shape and len(exp) > 0, (g.shape, exp.shape)
        bad = fit_mismatch_rows(g, exp)
        assert bad.size == 0, (bad.size, fit_errors(g, exp))
