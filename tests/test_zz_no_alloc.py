"""pt_fitting_noalloc (Fitting without allocation, its work space a fixed device array) against the
Fitting entry point: identical rows, on two calls in a row (the second with fewer positions), in the
reference-order FP64 fit and the fast fit, and over-size calls refused."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu
FIT_PARAMS = (7.0, 1.3, 1.3, 0.0)


@pytest.fixture(scope="module")
def stack(pt_gpu, oracle_mod):
    import torch
    d = torch.empty((24, 256, 256), dtype=torch.uint16, device="cuda")
    pt_gpu.synth_device(d, seed=0x5EED0777, emitters=60)
    st = d.cpu().numpy()
    thr = oracle_mod.autothreshold(st[0])
    det = pt_gpu.localization(st, thr, 0.0, 0, FIT_PARAMS)
    pos = np.column_stack([det[:, 0], det[:, 1], det[:, 4], det[:, 5], det[:, 15], det[:, 16]])
    # Input order not plane-major (the per-plane index follows input order), and one position whose
    # plane is outside the stack (failure row, counted in a bucket of its own).
    rng = np.random.default_rng(3)
    pos = pos[rng.permutation(len(pos))]
    pos[5, 1] = 99.0
    return d, st, np.ascontiguousarray(pos)


@pytest.mark.parametrize("fp64", [1, 0])
def test_noalloc_matches_fitting_two_calls(pt_gpu, stack, fp64):
    import torch
    d, st, pos = stack
    assert len(pos) > 1000
    pt_gpu.set_option("fit_fp64", fp64)
    try:
        for n in (len(pos), len(pos) // 3):
            exp = pt_gpu.fitting(st, pos[:n], 2, FIT_PARAMS)
            d_pos = torch.from_numpy(pos[:n]).cuda()
            out = torch.full((n * 18 + 7,), -5.0, dtype=torch.float64, device="cuda")
            rows = pt_gpu.fitting_noalloc(d, d_pos, out, 2, FIT_PARAMS)
            got = out[: n * 18].view(-1, 18).cpu().numpy()
            assert rows == n
            assert np.array_equal(got, exp, equal_nan=True)
            assert bool((out[n * 18:] == -5.0).all())
    finally:
        pt_gpu.set_option("fit_fp64", 1)


def test_noalloc_refuses_oversize_calls(pt_gpu, stack, capfd):
    import torch
    d, _, _ = stack
    big = torch.zeros(((1 << 17) + 1, 6), dtype=torch.float64, device="cuda")
    big[:, 1] = 1.0
    out = torch.empty(18, dtype=torch.float64, device="cuda")
    assert pt_gpu.fitting_noalloc(d, big, out, 2, FIT_PARAMS) == 0
    assert "at most 131072 positions" in capfd.readouterr().err
    pos = torch.zeros((4, 6), dtype=torch.float64, device="cuda")
    assert pt_gpu.fitting_noalloc(d, pos, out, 5, FIT_PARAMS) == 0
    assert "modes 1-4" in capfd.readouterr().err
