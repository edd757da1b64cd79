"""The fast fit (FP32 persistent kernel with the FP64 hand-over, fit mode 4 in the FP64 group kernel)
against the FP64 CPU oracle, at the sizes where its rare paths show: every row within the fit tolerance
(positions 1e-3 px; intensity, amplitude, offset, sigma 1e-4 relative; theta 1e-4 rad)."""
import os
import subprocess
import sys

import numpy as np
import pytest

from helpers import fit_errors, fit_mismatch_rows, fit_mismatches

pytestmark = pytest.mark.gpu
FIT_PARAMS = (7.0, 1.3, 1.3, 0.0)
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(autouse=True)
def fast_fit(pt_gpu):
    pt_gpu.set_option("fit_fp64", 0)
    pt_gpu.set_option("max_iter", 75)
    pt_gpu.set_option("fit_handover", 16)
    yield
    pt_gpu.set_option("fit_fp64", 1)
    pt_gpu.set_option("max_iter", 75)
    pt_gpu.set_option("fit_handover", 16)


@pytest.fixture(scope="module")
def c2_frames(pt_gpu, oracle_mod):
    """The first 200 frames of the 512^2 benchmark stack (seed 0x5EED0002, 200 emitters per frame)."""
    import torch
    d = torch.empty((200, 512, 512), dtype=torch.uint16, device="cuda")
    pt_gpu.synth_device(d, seed=0x5EED0002, emitters=200)
    st = d.cpu().numpy()
    return d, st, oracle_mod.autothreshold(st[0])


def test_fast_fit_200_frames_512(pt_gpu, oracle_mod, c2_frames):
    """200 frames of 512^2 (about 39 000 ROIs), fit mode 2, 7 x 7, at most 75 iterations: every row
    within the tolerance of the oracle, and the hand-over takes a few percent of the ROIs."""
    import torch
    d, st, thr = c2_frames
    out = torch.empty(200 * 1000 * 18, dtype=torch.float64, device="cuda")
    n = pt_gpu.localization_device(d, out, thr, 0.0, 2, FIT_PARAMS, 1000)
    handed = pt_gpu.last_stats()["fit_handed_over"]
    got = out[: n * 18].view(-1, 18).cpu().numpy()
    exp = oracle_mod.localization(st, thr, 0.0, 2, FIT_PARAMS, output_size=200 * 1000 * 18)
    assert got.shape == exp.shape and len(exp) > 30000
    bad = fit_mismatch_rows(got, exp)
    assert bad.size == 0, (bad.size, got[bad[:3]], fit_errors(got, exp))
    assert 0 < handed < 0.1 * len(exp), handed


@pytest.mark.parametrize("fit", [1, 2, 3, 4])
def test_fast_fit_iteration_cap_20(pt_gpu, oracle_mod, c2_frames, fit):
    """The 20-iteration cap of the benchmark configurations, every mode."""
    _, st, thr = c2_frames
    sub = np.ascontiguousarray(st[:40])
    pt_gpu.set_option("max_iter", 20)
    got = pt_gpu.localization(sub, thr, 0.0, fit, FIT_PARAMS)
    exp = oracle_mod.localization(sub, thr, 0.0, fit, FIT_PARAMS, max_iter=20)
    assert got.shape == exp.shape and len(exp) > 5000
    bad = fit_mismatch_rows(got, exp)
    assert bad.size == 0, (bad.size, fit_errors(got, exp))


@pytest.mark.parametrize("roi", [9, 11, 17])
@pytest.mark.parametrize("fit", [1, 2, 3, 4])
def test_fast_fit_roi_sides(pt_gpu, oracle_mod, c2_frames, roi, fit):
    """ROI side 9 (compiled kernels), 11 (side read at run time, staged in LDS) and 17 (read from the
    frame in every pass), in every mode."""
    _, st, thr = c2_frames
    sub = np.ascontiguousarray(st[:6])
    fp = (float(roi), 1.3, 1.3, 0.0)
    got = pt_gpu.localization(sub, thr, 0.0, fit, fp)
    exp = oracle_mod.localization(sub, thr, 0.0, fit, fp)
    assert got.shape == exp.shape and len(exp) > 800
    bad = fit_mismatches(got, exp)
    assert bad == [], (len({b[0] for b in bad}), bad[:4], fit_errors(got, exp))


@pytest.mark.parametrize("fit", [1, 2, 3])
def test_pure_fp32_fit_errors(pt_gpu, oracle_mod, c2_frames, fit):
    """Without the hand-over every ROI stays in FP32 (the precision extension): its rows are compared
    with the oracle and the error bound it keeps on this sample is stated here."""
    _, st, thr = c2_frames
    sub = np.ascontiguousarray(st[:40])
    pt_gpu.set_option("fit_handover", 0)
    got = pt_gpu.localization(sub, thr, 0.0, fit, FIT_PARAMS)
    exp = oracle_mod.localization(sub, thr, 0.0, fit, FIT_PARAMS)
    assert got.shape == exp.shape
    bad = fit_mismatch_rows(got, exp)
    # Out-of-tolerance rows come from fits that run to the iteration cap or meet a non-positive gain
    # (rounding-dominated, flat chi^2 valleys): about 1 in 10^4 rows in mode 2.
    assert bad.size <= max(3, len(exp) // 1000), (bad.size, fit_errors(got, exp))


def test_fitting_first_call_fast_fit_fresh_process(pt_gpu, oracle_mod, ref_vectors):
    """Fitting as the first call of a process with the fast fit selected (no Localization before it to
    allocate the fit's work space)."""
    st = ref_vectors["dense"]
    pos = ref_vectors["dense_positions"]
    code = (
        "import sys, numpy as np; sys.path.insert(0, %r); sys.path.insert(0, %r)\n"
        "import palmtracer_amd as pt\n"
        "pt.set_option('fit_fp64', 0)\n"
        "st = np.load(sys.argv[1]); pos = np.load(sys.argv[2])\n"
        "rows = pt.fitting(st, pos, 2, (7.0, 1.3, 1.3, 0.0))\n"
        "np.save(sys.argv[3], rows)\n" % (os.path.join(ROOT, "palmtracer-cpp_amd"), os.path.join(ROOT, "tests")))
    import tempfile
    with tempfile.TemporaryDirectory() as td:
        a, b, c = (os.path.join(td, n) for n in ("st.npy", "pos.npy", "rows.npy"))
        np.save(a, st)
        np.save(b, pos)
        r = subprocess.run([sys.executable, "-c", code, a, b, c], capture_output=True, text=True, timeout=300)
        assert r.returncode == 0, r.stderr
        got = np.load(c)
    exp = oracle_mod.fitting(st, pos, 2, (7.0, 1.3, 1.3, 0.0))
    assert got.shape == exp.shape and got.shape[0] == pos.shape[0]
    assert fit_mismatches(got, exp) == []


def test_fast_fit_matches_fp64_reference_order(pt_gpu, c2_frames):
    """Fast fit and the reference-order FP64 fit on the same frames: equal detection columns and fitted
    columns within the tolerance (both are checked against the oracle above; this pins that they are
    exchangeable through one option)."""
    _, st, thr = c2_frames
    sub = np.ascontiguousarray(st[:10])
    fast = pt_gpu.localization(sub, thr, 0.0, 2, FIT_PARAMS)
    pt_gpu.set_option("fit_fp64", 1)
    ref = pt_gpu.localization(sub, thr, 0.0, 2, FIT_PARAMS)
    assert fast.shape == ref.shape
    assert fit_mismatch_rows(fast, ref).size == 0


# Dense frames of the served and batched workloads (5000 emitters per frame): ROIs holding noise
# spikes or tangled spots, with fitted widths at their lower bound and barely any correlation, are
# decided in FP64 (a width at a bound, a squared correlation below 0.05).
DENSE = [(1024, 0x5EED1002), (2048, 0x5EED1003)]


@pytest.mark.parametrize("side,seed", DENSE)
def test_fast_fit_dense_frames(pt_gpu, oracle_mod, side, seed):
    """Ten dense frames at the 20-iteration cap: every fast-fit row within the tolerance of the oracle."""
    import torch
    d = torch.empty((10, side, side), dtype=torch.uint16, device="cuda")
    pt_gpu.synth_device(d, seed=seed, emitters=5000)
    st = d.cpu().numpy()
    thr = oracle_mod.autothreshold(st[0])
    pt_gpu.set_option("max_iter", 20)
    got = pt_gpu.localization(st, thr, 0.0, 2, FIT_PARAMS)
    handed = pt_gpu.last_stats()["fit_handed_over"]
    exp = oracle_mod.localization(st, thr, 0.0, 2, FIT_PARAMS, max_iter=20)
    assert got.shape == exp.shape and len(exp) > 40000
    bad = fit_mismatch_rows(got, exp)
    assert bad.size == 0, (bad.size, got[bad[:3]], fit_errors(got, exp))
    assert 0 < handed < 0.2 * len(exp), handed
