"""Astigmatic 3D localization on sCMOS-like 2048^2 frames: the 9 x 9 fit with free sigma_x / sigma_y
(fit mode 3, lib/fit/Fitting.cpp:23,147-152, lib/fit/GaussianFit.cpp:187-191) against the FP64
oracle, then z from the fitted widths by Astigmatism3DEstimation (lib/fit/AstigFit.cpp:36-72).

Frames: per-pixel offset ~ N(100, 3) ADU and gain ~ N(2, 0.1) (fixed per seed), 2000 emitters per
frame, z ~ U[-600, 600] nm, sigma_{x,y} = 1.2 * sqrt(1 + ((z -/+ 250) / 400)^2) px."""
import numpy as np
import pytest

from helpers import AMP, SX, SY, fit_errors, fit_mismatch_rows

pytestmark = pytest.mark.gpu
ROI = 9
FIT_PARAMS = (float(ROI), 1.4, 1.4, 0.0)
ASTIG = dict(astig=1, s0=1.2, c=250.0, d=400.0, zmin=-600.0, zmax=600.0, scmos=1, sd_offset=3.0, sd_gain=0.1)
PIXEL_NM = 100.0
# Calibration model of the generator in the form Astigmatism3DCalibration writes it: per axis
# [z0, w, c3, c4, A] with sigma = A * sqrt(1 + u^2 + c3 u^3 + c4 u^4) / pixel size, u = (z - z0) / w.
MODEL = [250.0, 400.0, 0.0, 0.0, 1.2 * PIXEL_NM, -250.0, 400.0, 0.0, 0.0, 1.2 * PIXEL_NM]
Z_MAX = 600.0


@pytest.fixture(scope="module")
def astig_frames(pt_gpu, oracle_mod):
    import torch
    d = torch.empty((2, 2048, 2048), dtype=torch.uint16, device="cuda")
    pt_gpu.synth_device(d, seed=0x5EED0003, emitters=2000, **ASTIG)
    st = d.cpu().numpy()
    thr = oracle_mod.autothreshold(st[0], (float(ROI),))
    truth = torch.empty((2, 2000, 5), dtype=torch.float32, device="cuda")
    pt_gpu.synth_emitters_device(truth, seed=0x5EED0003, height=2048, width=2048, **ASTIG)
    exp = oracle_mod.localization(st, thr, 0.0, 3, FIT_PARAMS)
    return st, thr, exp, truth.cpu().numpy()


@pytest.fixture(autouse=True)
def restore(pt_gpu):
    yield
    pt_gpu.set_option("fit_fp64", 1)


def test_scmos_frames_are_astigmatic(astig_frames):
    st, thr, exp, truth = astig_frames
    assert st.min() >= 80 and st.mean() > 140
    sx, sy = truth[..., 2], truth[..., 3]
    assert sx.min() >= 1.2 - 1e-6 and sy.min() >= 1.2 - 1e-6 and sx.max() > 2.5 and sy.max() > 2.5
    assert len(exp) > 3000


@pytest.mark.parametrize("fp64", [1, 0])
def test_astig_fit_9x9_matches_oracle(pt_gpu, astig_frames, fp64):
    """Every row of the reference-order FP64 fit and of the fast fit within the fit tolerance."""
    st, thr, exp, _ = astig_frames
    pt_gpu.set_option("fit_fp64", fp64)
    got = pt_gpu.localization(st, thr, 0.0, 3, FIT_PARAMS)
    assert got.shape == exp.shape
    bad = fit_mismatch_rows(got, exp)
    assert bad.size == 0, (bad.size, fit_errors(got, exp))
    assert (got[:, AMP] > 0).mean() > 0.8


def test_astig_z_chain_matches_oracle(pt_gpu, oracle_mod, astig_frames):
    """z from the fast fit's widths (GPU Astigmatism3DEstimation) against z from the oracle's rows
    (CPU estimation): within 2 nm, Brent's tolerance (1e-3 |z| + 1e-12) on the axial search; and
    the GPU estimation equals the oracle's on the same widths bit for bit."""
    st, thr, exp, _ = astig_frames
    pt_gpu.set_option("fit_fp64", 0)
    got = pt_gpu.localization(st, thr, 0.0, 3, FIT_PARAMS)
    from helpers import pair_rows_by_plane
    j = pair_rows_by_plane(got, exp)
    ok = (j >= 0) & (got[:, AMP] > 0)
    ok &= exp[np.where(j >= 0, j, 0), AMP] > 0
    g, e = got[ok], exp[j[ok]]
    z_gpu = pt_gpu.astig_estimate(g[:, [SX, SY]], PIXEL_NM, MODEL, Z_MAX)
    z_ora = oracle_mod.astig_estimate(e[:, [SX, SY]], PIXEL_NM, MODEL, Z_MAX)
    assert np.array_equal(z_gpu, oracle_mod.astig_estimate(g[:, [SX, SY]], PIXEL_NM, MODEL, Z_MAX))
    dz = np.abs(z_gpu[:, 0] - z_ora[:, 0])
    assert dz.max() <= 2.0, (dz.max(), np.argmax(dz))
    assert np.all(np.abs(z_gpu[:, 0]) <= Z_MAX) and np.all((z_gpu[:, 1] >= 0) & (z_gpu[:, 1] <= 1))


def test_astig_z_recovers_the_emitters(pt_gpu, astig_frames):
    """Fitted widths through the calibration model recover the z of isolated bright emitters: the
    z of the true widths (same estimator) against the z of the fitted rows nearest to them."""
    st, thr, _, truth = astig_frames
    pt_gpu.set_option("fit_fp64", 0)
    got = pt_gpu.localization(st, thr, 0.0, 3, FIT_PARAMS)
    errs = []
    for f in range(st.shape[0]):
        t = truth[f]
        rows = got[(got[:, 1] == f + 1) & (got[:, AMP] > 0)]
        # Emitters with no other emitter within 8 px and more than 2000 photons.
        dd = np.hypot(t[:, None, 0] - t[None, :, 0], t[:, None, 1] - t[None, :, 1])
        np.fill_diagonal(dd, 1e9)
        iso = (dd.min(1) > 8) & (t[:, 4] > 2000) & (t[:, 0] > 6) & (t[:, 0] < 2042) & (t[:, 1] > 6) & (t[:, 1] < 2042)
        # Pixel centres are at integer coordinates, as in the model the fit evaluates.
        tx, ty = t[iso, 0], t[iso, 1]
        d2 = (rows[None, :, 4] - tx[:, None]) ** 2 + (rows[None, :, 5] - ty[:, None]) ** 2
        k = d2.argmin(1)
        near = d2[np.arange(len(k)), k] < 0.25
        # The fit's point-sampled Gaussian of a pixel-integrated spot has width sqrt(sigma^2 + 1/12).
        w_true = np.sqrt(t[iso][near][:, 2:4].astype(np.float64) ** 2 + 1.0 / 12.0)
        z_true = pt_gpu.astig_estimate(w_true, PIXEL_NM, MODEL, Z_MAX)[:, 0]
        z_fit = pt_gpu.astig_estimate(rows[k[near]][:, [SX, SY]], PIXEL_NM, MODEL, Z_MAX)[:, 0]
        errs.append(z_fit - z_true)
    e = np.concatenate(errs)
    assert len(e) > 200
    assert np.median(np.abs(e)) < 60.0, np.median(np.abs(e))
