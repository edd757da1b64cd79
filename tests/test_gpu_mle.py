"""Poisson maximum-likelihood objective of the Gaussian fits (option fit_objective = 1; an extension
beside the reference's least squares, lib/fit/GaussianFit.cpp:48-295): every fitted row against the
oracle's likelihood maximum (scipy L-BFGS-B on the same objective, an optimiser independent of the
library's Fisher-scoring LM) at the same ROI, and the spread over Poisson realisations at the
Cramer-Rao bound (the fit is efficient)."""
import numpy as np
import pytest

from helpers import pair_rows_by_plane

pytestmark = pytest.mark.gpu
CAMERA = (2.0, 100.0, 0.0)


@pytest.fixture(autouse=True)
def objective(pt_gpu):
    pt_gpu.set_option("fit_objective", 1)
    pt_gpu.set_option("camera_gain", CAMERA[0])
    pt_gpu.set_option("camera_offset", CAMERA[1])
    pt_gpu.set_option("camera_read_var", CAMERA[2])
    yield
    pt_gpu.set_option("fit_objective", 0)
    pt_gpu.set_option("camera_gain", 1.0)
    pt_gpu.set_option("camera_offset", 0.0)


@pytest.mark.parametrize("fit,fp", [(2, (7, 1.3, 1.3, 0.0)), (1, (7, 1.3, 1.3, 0.0)), (3, (9, 1.4, 1.4, 0.0))])
def test_mle_rows_match_oracle(pt_gpu, oracle_mod, fit, fp):
    import torch
    d = torch.empty((3, 256, 256), dtype=torch.uint16, device="cuda")
    pt_gpu.synth_device(d, seed=0x5EED8001 + fit, emitters=50)
    st = d.cpu().numpy()
    thr = oracle_mod.autothreshold(st[0])
    regions = oracle_mod.localization(st, thr, 0.0, 0, fp)
    got = pt_gpu.localization(st, thr, 0.0, fit, fp)
    assert got.shape == regions.shape and len(got) > 100
    exp = np.full_like(got, np.nan)
    exp[:, :6] = regions[:, :6]
    for k, r in enumerate(regions):
        q = oracle_mod.mle_fit(st[int(r[1]) - 1], fit, fp, r[4], r[5], CAMERA)
        if q is not None:
            exp[k, [14, 13, 4, 5, 8, 9, 10]] = q
        else:
            exp[k, 13] = -1.0
    j = pair_rows_by_plane(got, exp)
    assert (j >= 0).all() and len(np.unique(j)) == len(j)
    e = exp[j]
    # ROIs leaving the frame fail on both sides (the reference's border rule).
    border = e[:, 13] == -1
    assert (got[border, 13] == -1).all()
    ok = ~border
    assert (got[ok, 13] > 0).all()
    g, e = got[ok], e[ok]
    assert np.abs(g[:, 4] - e[:, 4]).max() <= 1e-3 and np.abs(g[:, 5] - e[:, 5]).max() <= 1e-3
    for c in (13, 14, 8, 9):
        rel = np.abs(g[:, c] - e[:, c]) / np.abs(e[:, c])
        assert rel.max() <= 1e-3, (c, rel.max())


def test_mle_spread_at_crlb(pt_gpu, oracle_mod):
    """600 frames of one emitter drawn from the fit's own model: the spread of the fitted positions
    and widths equals the bound (pt_crlb_device at the true parameters) within sampling error."""
    import torch
    H = W = 21
    X0, Y0, amp, off = 10.3, 10.6, 800.0, 140.0
    yy, xx = np.mgrid[0:H, 0:W].astype(float)
    lam = (oracle_mod.gauss_model(2, np.array([off, amp, X0, Y0, 1.3]), xx - X0, yy - Y0, 1.3, 0.0) - 100.0) / 2.0
    rng = np.random.default_rng(31)
    st = np.ascontiguousarray((100 + 2 * rng.poisson(lam, size=(600, H, W))).astype(np.uint16))
    pos = np.zeros((600, 6))
    pos[:, 0] = np.arange(1, 601)
    pos[:, 1] = np.arange(1, 601)
    pos[:, 2], pos[:, 3] = 10.0, 11.0
    rows = pt_gpu.fitting(st, pos, 2, (7, 1.3, 1.3, 0.0))
    assert (rows[:, 13] > 0).all()
    truth = np.zeros((1, 18))
    truth[0, [4, 5, 13, 14, 8, 9]] = X0, Y0, amp, off, 1.3, 1.3
    c = pt_gpu.crlb_device(torch.from_numpy(truth).cuda(), H, W, 2, (7, 1.3, 1.3, 0.0),
                           torch.empty((1, 8), dtype=torch.float64, device="cuda"), CAMERA).cpu().numpy()[0]
    for k, col in ((0, 4), (1, 5), (2, 8), (5, 13), (6, 14)):
        ratio = rows[:, col].std() / c[k]
        assert 0.9 <= ratio <= 1.12, (k, ratio)
    assert abs(rows[:, 4].mean() - X0) < 4 * c[0] / np.sqrt(600)
