"""Exact frame flood of dense frames (Segmentation2D.cpp:336-385) through the grid rank flood: every
frame's labels equal the oracle's, and the frames are resolved by the grid rank flood itself (not
handed to the single-workgroup floods). Frames: the generator's 2D spots at up to 5 emitters per
um^2 at 100 nm pixels (capped at 5000 per frame), the densities of batched and served workloads."""
import numpy as np
import pytest

from helpers import filter_small, rank_labels

pytestmark = pytest.mark.gpu
FIT_PARAMS = (7.0, 1.3, 1.3, 0.0)
DENSE = [(2048, 5000), (1024, 5000), (512, 5000), (256, 3277), (128, 819)]


@pytest.fixture(scope="module")
def dense_frames(pt_gpu, oracle_mod):
    import torch
    out = {}
    for side, em in DENSE:
        d = torch.empty((2, side, side), dtype=torch.uint16, device="cuda")
        pt_gpu.synth_device(d, seed=0x5EED4000 + side + em, emitters=em)
        st = d.cpu().numpy()
        out[side] = (st, oracle_mod.autothreshold(st[0]))
    return out


@pytest.mark.parametrize("side", [s for s, _ in DENSE])
def test_grid_flood_labels(pt_gpu, oracle_mod, dense_frames, side):
    st, thr = dense_frames[side]
    pt_gpu.set_option("force_sequential_watershed", 1)
    try:
        for f in range(st.shape[0]):
            w1 = oracle_mod.atrous(st[f])[1]
            exp = oracle_mod.watershed_labels(w1, thr)
            got = pt_gpu.debug_labels(st[f], thr, 0.0)
            stats = pt_gpu.last_stats()
            assert stats["status"] == 0
            assert np.array_equal(rank_labels(got), exp), (side, f, stats)
            assert stats["grid_frames"] == 1, stats
    finally:
        pt_gpu.set_option("force_sequential_watershed", 0)


@pytest.mark.parametrize("side", [2048, 512])
def test_grid_flood_second_watershed(pt_gpu, oracle_mod, dense_frames, side):
    """Basin fusion (watershed ratio 10): both watersheds through the grid rank flood."""
    st, thr = dense_frames[side]
    pt_gpu.set_option("force_sequential_watershed", 1)
    try:
        w1 = oracle_mod.atrous(st[0])[1]
        exp, _ = oracle_mod.segment(w1, thr, 10.0)
        got = pt_gpu.debug_labels(st[0], thr, 10.0)
        stats = pt_gpu.last_stats()
    finally:
        pt_gpu.set_option("force_sequential_watershed", 0)
    assert stats["status"] == 0
    assert np.array_equal(filter_small(got), exp), stats
    assert stats["grid_frames"] >= 1, stats


@pytest.mark.parametrize("side", [2048, 1024, 128])
def test_dense_localization_rows(pt_gpu, oracle_mod, dense_frames, side):
    """Localization (regions, fit mode 0) of the dense frames equals the oracle's rows."""
    st, thr = dense_frames[side]
    got = pt_gpu.localization(st, thr, 0.0, 0, FIT_PARAMS)
    exp = oracle_mod.localization(st, thr, 0.0, 0, FIT_PARAMS)
    assert got.shape == exp.shape and np.array_equal(got, exp)


def mirrored_frame(seed, h=48, half=24, spots=4):
    """A frame mirrored about its vertical centre line, spots near the centre: every potential has
    an equal twin, and spots joined across the centre make equal keys whose order can decide the
    labels there (background 20 photons, sigma 1.3 px, ADU = 100 + 2 Poisson)."""
    from scipy.special import erf
    rng = np.random.default_rng(seed)
    lam = np.full((h, half), 20.0)
    s2 = np.sqrt(2.0) * 1.3
    px, py = np.arange(half), np.arange(h)
    for _ in range(spots):
        x, y, k = rng.uniform(half - 4.0, half - 0.6), rng.uniform(4, h - 4), rng.uniform(800, 3000)
        gx = 0.5 * (erf((px + 0.5 - x) / s2) - erf((px - 0.5 - x) / s2))
        gy = 0.5 * (erf((py + 0.5 - y) / s2) - erf((py - 0.5 - y) / s2))
        lam += k * np.outer(gy, gx)
    left = np.minimum(100 + 2 * rng.poisson(lam), 65535).astype(np.uint16)
    return np.ascontiguousarray(np.concatenate([left, left[:, ::-1]], axis=1))


# (seed, the order of equal keys can change a label): solved on the CPU under many orders.
MIRRORED = [(7, True), (1, False), (2, False), (4, False), (18, False)]


@pytest.mark.parametrize("seed,order_matters", MIRRORED)
def test_mirrored_frames(pt_gpu, oracle_mod, seed, order_matters):
    """Equal keys everywhere. The grid rank flood accepts a frame only when no order of the equal keys
    can change a label, and hands the others to the single-workgroup floods; labels equal the
    oracle's either way."""
    frame = mirrored_frame(seed)
    thr = oracle_mod.autothreshold(frame)
    exp = oracle_mod.watershed_labels(oracle_mod.atrous(frame)[1], thr)
    pt_gpu.set_option("force_sequential_watershed", 1)
    try:
        got = pt_gpu.debug_labels(frame, thr, 0.0)
        stats = pt_gpu.last_stats()
    finally:
        pt_gpu.set_option("force_sequential_watershed", 0)
    assert np.array_equal(rank_labels(got), exp), stats
    if order_matters:
        assert stats["grid_frames"] == 0 and stats["grid_rejected"] == 1, stats
    else:
        assert stats["grid_frames"] == 1 and stats["grid_equal_keys"] == 1, stats
