"""Concurrent calls of the C entry points from several host threads run independently and return
exactly what the same calls return one at a time (Tests_CPU/test_fit_Fitting.cpp:337-365 runs one
fitter from 4 OpenMP threads and requires every ROI's row to equal the serial one). ctypes releases
the interpreter lock during a foreign call, so the threads' calls overlap inside the library: more
threads than the library has engines also exercises the wait for a free one."""
import threading

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
FIT_PARAMS = (7.0, 1.3, 1.3, 0.0)
SHAPES = [(6, 256, 60), (4, 512, 200), (10, 128, 20), (3, 384, 120)]   # frames, side, emitters


@pytest.fixture(scope="module")
def stacks(pt_gpu):
    import torch
    out = []
    for k, (n, side, emitters) in enumerate(SHAPES):
        d = torch.empty((n, side, side), dtype=torch.uint16, device="cuda")
        pt_gpu.synth_device(d, seed=0x5EED7001 + k, emitters=emitters)
        torch.cuda.synchronize()
        st = d.cpu().numpy()
        thr = pt_gpu.autothreshold(st[0])
        pos = pt_gpu.localization(st, thr, 0.0, 0, FIT_PARAMS)[:, [0, 1, 4, 5, 15, 16]].copy()
        out.append((st, d, thr, pos))
    return out


def _calls(pt_gpu, stacks):
    import torch
    calls = []
    for k, (st, d, thr, pos) in enumerate(stacks):
        mode = 1 + k % 4
        calls.append(("Localization", k, lambda st=st, thr=thr: pt_gpu.localization(st, thr, 0.0, 2, FIT_PARAMS)))
        calls.append(("Fitting", k, lambda st=st, pos=pos, mode=mode: pt_gpu.fitting(st, pos, mode, FIT_PARAMS)))
        calls.append(("Wavelett", k, lambda st=st: pt_gpu.wavelett(st, 2)))
        calls.append(("AutoThreshold", k, lambda st=st: np.array([pt_gpu.autothreshold(st[-1])])))

        def device(d=d, thr=thr, n=st.shape[0], e=SHAPES[k][2]):
            mpp = 2 * e + 64
            out = torch.empty(n * mpp * 18, dtype=torch.float64, device="cuda")
            m = pt_gpu.localization_device(d, out, thr, 0.0, 2, FIT_PARAMS, mpp)
            return out[:m * 18].reshape(-1, 18).cpu().numpy()
        calls.append(("pt_localization_device", k, device))
    return calls


def test_concurrent_calls_match_serial(pt_gpu, stacks):
    calls = _calls(pt_gpu, stacks)
    serial = [f() for _, _, f in calls]
    assert all(r.size for r in serial)
    threads, rounds = 12, 2
    got = [[None] * len(calls) for _ in range(threads)]
    errors = []
    start = threading.Barrier(threads)

    def worker(t):
        try:
            order = np.random.default_rng(100 + t).permutation(len(calls) * rounds) % len(calls)
            start.wait()
            for i in order:
                r = calls[i][2]()
                if got[t][i] is not None and not np.array_equal(r, got[t][i], equal_nan=True):
                    errors.append((t, calls[i][0], calls[i][1], "differs between rounds"))
                got[t][i] = r
        except Exception as ex:   # reported below, with the thread
            errors.append((t, repr(ex)))

    pool = [threading.Thread(target=worker, args=(t,)) for t in range(threads)]
    for th in pool:
        th.start()
    for th in pool:
        th.join(timeout=110)
    assert not any(th.is_alive() for th in pool), "a concurrent call did not return"
    assert not errors, errors[:5]
    for t in range(threads):
        for i, (name, k, _) in enumerate(calls):
            assert np.array_equal(got[t][i], serial[i], equal_nan=True), (t, name, k)


def test_options_reach_every_engine(pt_gpu, stacks):
    """An option set between calls applies to whichever engine serves the next call: of four
    simultaneous calls at most one runs on the process-wide engine, and every one must see max_iter."""
    st, _, thr, _ = stacks[1]
    pt_gpu.set_option("max_iter", 1)
    try:
        one = pt_gpu.localization(st, thr, 0.0, 2, FIT_PARAMS)
        res = [None] * 4
        barrier = threading.Barrier(4)

        def worker(t):
            barrier.wait()
            res[t] = pt_gpu.localization(st, thr, 0.0, 2, FIT_PARAMS)

        pool = [threading.Thread(target=worker, args=(t,)) for t in range(4)]
        for th in pool:
            th.start()
        for th in pool:
            th.join(timeout=110)
        assert not any(th.is_alive() for th in pool)
        for r in res:
            assert np.array_equal(r, one, equal_nan=True)
    finally:
        pt_gpu.set_option("max_iter", 75)   # lib/fit/Fitting.hpp:80
    full = pt_gpu.localization(st, thr, 0.0, 2, FIT_PARAMS)
    assert not np.array_equal(full, one, equal_nan=True)


def test_counters_follow_the_last_counting_call(pt_gpu, stacks):
    """pt_last_stats reports the last call that computes counters, whichever engine ran it: a
    Localization's rows, then a Fitting's frames and positions; Wavelett and the astigmatism
    estimate leave them as they were."""
    st, _, thr, pos = stacks[1]
    rows = pt_gpu.localization(st, thr, 0.0, 2, FIT_PARAMS)
    s = pt_gpu.last_stats()
    assert s["rows"] == len(rows) and s["frames"] == st.shape[0]
    pt_gpu.wavelett(st, 2)
    pt_gpu.astig_estimate(np.full((4, 2), 1.3), 100.0, [0.0, 1.0, 0.0, 0.0, 1.0] * 2, 600.0)
    assert pt_gpu.last_stats() == s
    sub = pos[: len(pos) // 2]
    pt_gpu.fitting(st, sub, 2, FIT_PARAMS)
    s2 = pt_gpu.last_stats()
    assert s2["rows"] == len(sub) and s2["frames"] == st.shape[0] and s2["flagged_frames"] == 0
