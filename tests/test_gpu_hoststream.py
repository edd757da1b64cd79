"""Localization / Fitting over host buffers, the stack streamed through the device in pieces
(CPU/main.cpp:112-121, 141-148; lib/Palm.cpp:24-43 for the global Id and plane numbering): rows equal
the single-piece call and the oracle whatever the piece size."""
import numpy as np
import pytest

from helpers import fit_mismatch_rows

pytestmark = pytest.mark.gpu
FIT_PARAMS = (7.0, 1.3, 1.3, 0.0)


@pytest.fixture(scope="module")
def stack(pt_gpu, oracle_mod):
    import torch
    d = torch.empty((12, 256, 256), dtype=torch.uint16, device="cuda")
    pt_gpu.synth_device(d, seed=0x5EED6001, emitters=60)
    st = d.cpu().numpy()
    return st, oracle_mod.autothreshold(st[0])


@pytest.fixture(autouse=True)
def restore(pt_gpu):
    yield
    pt_gpu.set_option("stream_pixels", 0)
    pt_gpu.set_option("host_threads", 8)


@pytest.mark.parametrize("frames_per_piece", [1, 5, 12])
@pytest.mark.parametrize("fit", [0, 2])
def test_streamed_localization(pt_gpu, oracle_mod, stack, frames_per_piece, fit):
    st, thr = stack
    pt_gpu.set_option("stream_pixels", frames_per_piece * 256 * 256)
    got = pt_gpu.localization(st, thr, 0.0, fit, FIT_PARAMS)
    pieces = pt_gpu.stage_times()["stream_pieces"]
    assert pieces == -(-12 // frames_per_piece)
    pt_gpu.set_option("stream_pixels", 0)
    whole = pt_gpu.localization(st, thr, 0.0, fit, FIT_PARAMS)
    assert np.array_equal(got, whole)
    exp = oracle_mod.localization(st, thr, 0.0, fit, FIT_PARAMS)
    assert got.shape == exp.shape and got.shape[0] > 300
    if fit == 0:
        assert np.array_equal(got, exp)
    else:
        assert fit_mismatch_rows(got, exp).size == 0
    assert np.array_equal(got[:, 0], np.arange(1, got.shape[0] + 1))     # global Id over the pieces


def test_streamed_localization_capacity(pt_gpu, oracle_mod, stack):
    """A tight output: at most (outputSize / 18) / planes rows per plane in every piece."""
    st, thr = stack
    size = 18 * 12 * 9 + 5
    pt_gpu.set_option("stream_pixels", 5 * 256 * 256)
    got = pt_gpu.localization(st, thr, 0.0, 0, FIT_PARAMS, output_size=size)
    exp = oracle_mod.localization(st, thr, 0.0, 0, FIT_PARAMS, output_size=size)
    assert got.shape == exp.shape and np.array_equal(got, exp)
    assert np.bincount(got[:, 1].astype(int)).max() <= 9


@pytest.mark.parametrize("threads", [1, 8])
def test_streamed_fitting(pt_gpu, oracle_mod, stack, threads):
    """Positions of every plane, shuffled: each piece fits the positions of its planes, rows come back
    in input order with the caller's plane; one position outside the stack gets a failure row."""
    st, thr = stack
    rows = oracle_mod.localization(st, thr, 0.0, 0, FIT_PARAMS)
    pos = rows[:, [0, 1, 4, 5, 15, 16]].copy()
    rng = np.random.default_rng(3)
    pos = pos[rng.permutation(len(pos))]
    pos[7, 1] = 99.0
    pt_gpu.set_option("host_threads", threads)
    pt_gpu.set_option("stream_pixels", 3 * 256 * 256)
    got = pt_gpu.fitting(st, pos, 2, FIT_PARAMS)
    assert pt_gpu.stage_times()["stream_pieces"] == 4
    pt_gpu.set_option("stream_pixels", 0)
    whole = pt_gpu.fitting(st, pos, 2, FIT_PARAMS)
    assert np.array_equal(got, whole, equal_nan=True)
    assert got.shape == (len(pos), 18) and got[7, 1] == 99.0
    ok = np.ones(len(pos), bool)
    ok[7] = False
    exp = oracle_mod.fitting(st, pos[ok], 2, FIT_PARAMS)
    assert fit_mismatch_rows(got[ok], exp).size == 0
