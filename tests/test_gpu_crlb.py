"""pt_crlb_device (Gaussian fit modes 1-4) against the oracle's bounds (finite-difference Fisher
information, its own accuracy about 1e-7) on the rows of a Localization: equal to 2e-6 relative,
NaN in the same places."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu
CAMERA = (2.0, 100.0, 0.0)


@pytest.fixture(scope="module")
def stack(pt_gpu, oracle_mod):
    import torch
    d = torch.empty((4, 256, 256), dtype=torch.uint16, device="cuda")
    pt_gpu.synth_device(d, seed=0x5EED7001, emitters=80)
    st = d.cpu().numpy()
    return st, oracle_mod.autothreshold(st[0])


@pytest.mark.parametrize("fit,fp", [(1, (7, 1.3, 1.3, 0.0)), (2, (7, 1.3, 1.3, 0.0)), (3, (9, 1.4, 1.4, 0.0)),
                                    (4, (9, 1.4, 1.4, 0.2))])
def test_crlb_matches_oracle(pt_gpu, oracle_mod, stack, fit, fp):
    import torch
    st, thr = stack
    rows = pt_gpu.localization(st, thr, 0.0, fit, fp)
    assert len(rows) > 200
    d_rows = torch.from_numpy(rows).cuda()
    d_out = torch.empty((len(rows), 8), dtype=torch.float64, device="cuda")
    got = pt_gpu.crlb_device(d_rows, 256, 256, fit, fp, d_out, CAMERA).cpu().numpy()
    exp = oracle_mod.crlb(rows, 256, 256, fit, fp, CAMERA)
    assert np.array_equal(np.isnan(got), np.isnan(exp))
    ok = ~np.isnan(exp)
    assert ok[:, 0].mean() > 0.8
    np.testing.assert_allclose(got[ok], exp[ok], rtol=2e-6, atol=1e-12)
    assert (got[ok] > 0).all()
