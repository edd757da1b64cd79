"""Cramer-Rao bounds of the Gaussian fits: the oracle's bounds (finite-difference Fisher information)
against the spread of least-squares fits over Poisson realisations of the fit's own model. Least
squares is not efficient under Poisson noise, so its spread lies at or somewhat above the bound:
each ratio must lie in [0.93, 1.6] (600 realisations: about 3 % sampling error on a standard
deviation), which a bound off by a factor of sqrt(2) (e.g. the gain left out of the variance) fails."""
import numpy as np
import pytest

H = W = 21
CASES = [(1, (7, 1.3, 1.3, 0.0), None), (2, (7, 1.3, 1.3, 0.0), 1.3), (3, (9, 1.4, 1.4, 0.0), (1.5, 1.2))]


@pytest.mark.parametrize("fit,fp,sig", CASES)
def test_oracle_crlb_against_fit_spread(oracle_mod, fit, fp, sig):
    X0, Y0, amp, off = 10.3, 10.6, 800.0, 140.0
    p = [off, amp, X0, Y0]
    if fit == 2:
        p += [sig]
    if fit == 3:
        p += [np.sqrt(2) * sig[0], np.sqrt(2) * sig[1]]
    yy, xx = np.mgrid[0:H, 0:W].astype(float)
    lam = (oracle_mod.gauss_model(fit, np.array(p), xx - X0, yy - Y0, fp[1], fp[3]) - 100.0) / 2.0
    rng = np.random.default_rng(11 + fit)
    fits = np.array([oracle_mod.fit_region(fit, (100 + 2 * rng.poisson(lam)).astype(np.uint16), fp, 10.0, 11.0)
                     for _ in range(600)])
    assert (fits[:, 13] > 0).all()
    truth = np.zeros((1, 18))
    truth[0, [4, 5, 13, 14]] = X0, Y0, amp, off
    if fit == 2:
        truth[0, 8] = truth[0, 9] = sig
    if fit == 3:
        truth[0, 8], truth[0, 9] = sig
    c = oracle_mod.crlb(truth, H, W, fit, fp, (2.0, 100.0, 0.0))[0]
    cols = {0: 4, 1: 5, 5: 13, 6: 14, 7: 7}
    if fit >= 2:
        cols.update({2: 8, 3: 9})
    for k, col in cols.items():
        ratio = fits[:, col].std() / c[k]
        assert 0.93 <= ratio <= 1.6, (k, ratio, c[k])
    assert np.isnan(c[4])
    assert abs(fits[:, 4].mean() - X0) < 3 * c[0] / np.sqrt(600) * 1.6 + 2e-3


def test_oracle_crlb_failed_and_border_rows(oracle_mod):
    rows = np.zeros((3, 18))
    rows[:, [4, 5, 13, 14, 8, 9]] = [10.0, 10.0, 500.0, 120.0, 1.3, 1.3]
    rows[1, 13] = -1.0          # failed fit
    rows[2, 4] = 1.5            # ROI beyond the frame
    c = oracle_mod.crlb(rows, H, W, 2, (7, 1.3, 1.3, 0.0), (2.0, 100.0, 0.0))
    assert np.isfinite(c[0, [0, 1, 2, 3, 5, 6, 7]]).all()
    assert np.isnan(c[1]).all() and np.isnan(c[2]).all()


def test_oracle_mle_reaches_crlb(oracle_mod):
    """The Poisson maximum-likelihood fit is efficient: its spread over 400 realisations equals the
    bound within sampling error (band [0.88, 1.12]), which pins both the likelihood and the bound."""
    X0, Y0, amp, off = 10.3, 10.6, 800.0, 140.0
    p = np.array([off, amp, X0, Y0, 1.3])
    yy, xx = np.mgrid[0:H, 0:W].astype(float)
    lam = (oracle_mod.gauss_model(2, p, xx - X0, yy - Y0, 1.3, 0.0) - 100.0) / 2.0
    rng = np.random.default_rng(21)
    fits = np.array([oracle_mod.mle_fit((100 + 2 * rng.poisson(lam)).astype(np.uint16), 2, (7, 1.3, 1.3, 0.0), 10.0, 11.0,
                                        (2.0, 100.0, 0.0)) for _ in range(400)])
    truth = np.zeros((1, 18))
    truth[0, [4, 5, 13, 14, 8, 9]] = X0, Y0, amp, off, 1.3, 1.3
    c = oracle_mod.crlb(truth, H, W, 2, (7, 1.3, 1.3, 0.0), (2.0, 100.0, 0.0))[0]
    for k, col in ((0, 2), (1, 3), (2, 4), (5, 1), (6, 0)):
        ratio = fits[:, col].std() / c[k]
        assert 0.88 <= ratio <= 1.12, (k, ratio)
