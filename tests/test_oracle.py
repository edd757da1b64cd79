"""Pins the CPU oracle (oracle/oracle.c) against the reference's golden data, its known answers and
mathematics that fixes the result independently of the restatement. CPU only."""
import numpy as np
import pytest

from helpers import CIRC, ID, INDEX, INTENSITY, PLANE, QXY, SX, SY, THETA, X, Y, match_rows, rank_labels, theta_to_degrees

FIT_PARAMS_CSV = (7.0, 1.0, 2.0, 0.0)  # Tests_CPU/test_Palm.cpp:25


# ---------------------------------------------------------------- a trous -----------------------
def _b3_direct(img, step):
    """Separable [1,4,6,4,1]/16 with taps at multiples of `step`, clamp to edge (independent of oracle)."""
    h, w = img.shape
    taps = [(-2, 1 / 16), (-1, 1 / 4), (0, 3 / 8), (1, 1 / 4), (2, 1 / 16)]
    cols = np.arange(w)
    rows = np.arange(h)
    tmp = sum(c * img[:, np.clip(cols + k * step, 0, w - 1)] for k, c in taps)
    return sum(c * tmp[np.clip(rows + k * step, 0, h - 1), :] for k, c in taps)


@pytest.mark.parametrize("shape", [(1, 1), (9, 1), (1, 9), (13, 7), (128, 256)])
def test_atrous_invariants(oracle_mod, shape):
    rng = np.random.default_rng(3)
    img = rng.integers(0, 65536, shape, dtype=np.uint16)
    w0, w1, v2 = oracle_mod.atrous(img)
    f = img.astype(np.float64)
    # Reconstruction: W0 + W1 + V2 == image (Tests_CPU/test_loc_ATrous.cpp:27-39), exactly.
    assert np.array_equal(w0 + w1 + v2, f)
    # W0 == image - direct B3 convolution (Tests_CPU/test_loc_ATrous.cpp:42-62).
    v1 = _b3_direct(f, 1)
    assert np.array_equal(w0, f - v1)
    assert np.array_equal(v2, _b3_direct(v1, 2))
    # The planes are dyadic: W1 * 2^16 is an integer (exact arithmetic in FP64).
    assert np.array_equal(w1 * 65536, np.round(w1 * 65536))


def test_atrous_constant_image(oracle_mod):
    img = np.full((17, 23), 1234, np.uint16)
    w0, w1, v2 = oracle_mod.atrous(img)
    assert not w0.any() and not w1.any()
    assert np.array_equal(v2, np.full(img.shape, 1234.0))


def test_atrous_matches_compiled_reference(oracle_mod, ref_inputs):
    R = oracle_mod.ref_lib()
    if R is None:
        pytest.skip("compiled reference not built here")
    st = ref_inputs["stack"]
    for f in range(st.shape[0]):
        img = np.ascontiguousarray(st[f])
        h, w = img.shape
        ref = [np.empty((h, w)) for _ in range(3)]
        R.ref_atrous(img.ctypes.data, h, w, *[p.ctypes.data for p in ref])
        for a, b in zip(oracle_mod.atrous(img), ref):
            assert np.array_equal(a, b)


# ---------------------------------------------------------------- segmentation ------------------
def _potential(h, w, peaks):
    """Hand-built W1 potential (the reference's segmentation tests inject it directly)."""
    yy, xx = np.mgrid[0:h, 0:w]
    p = np.zeros((h, w))
    for (y, x, a, s) in peaks:
        p += a * np.exp(-((yy - y) ** 2 + (xx - x) ** 2) / (2 * s * s))
    return p


def test_segment_threshold_is_strict(oracle_mod):
    w1 = np.zeros((9, 9))
    w1[3:6, 3:6] = 5.0
    w1[4, 4] = 6.0
    _, reg = oracle_mod.segment(w1, 5.0)
    assert len(reg) == 0  # only the centre is > 5: surface 1 <= VOLUME_MIN
    _, reg = oracle_mod.segment(w1, 4.999)
    assert len(reg) == 1 and reg[0, 4] == 9


def test_segment_volume_min_boundary(oracle_mod):
    # Surface 4 is rejected, 5 kept (Tests_CPU/test_loc_Segmentation2D.cpp:186-199).
    for n, kept in ((4, 0), (5, 1)):
        w1 = np.zeros((8, 8))
        w1[2, 2:2 + n] = np.arange(1, n + 1)
        _, reg = oracle_mod.segment(w1, 0.5)
        assert len(reg) == kept


def test_segment_weighted_centroid(oracle_mod):
    w1 = np.zeros((10, 10))
    vals = {(4, 4): 10.0, (4, 5): 6.0, (5, 4): 3.0, (5, 5): 1.0, (3, 4): 2.0}
    for (y, x), v in vals.items():
        w1[y, x] = v
    _, reg = oracle_mod.segment(w1, 0.0)
    s = sum(vals.values())
    assert reg.shape[0] == 1
    assert reg[0, 0] == pytest.approx(sum(v * x for (y, x), v in vals.items()) / s, abs=1e-12)
    assert reg[0, 1] == pytest.approx(sum(v * y for (y, x), v in vals.items()) / s, abs=1e-12)
    assert reg[0, 3] == s and reg[0, 4] == 5


def test_segment_two_peaks_merge_or_separate(oracle_mod):
    # Tests_CPU/test_loc_Segmentation2D.cpp:236-265: ratio 0 keeps two basins, ratio 10 merges them.
    w1 = _potential(20, 30, [(10, 9, 10.0, 2.0), (10, 19, 8.0, 2.0)])
    _, sep = oracle_mod.segment(w1, 0.5, 0.0)
    _, mer = oracle_mod.segment(w1, 0.5, 10.0)
    assert len(sep) == 2 and len(mer) == 1


def test_segment_many_regions(oracle_mod):
    # More than 65535 regions need u32 labels (Tests_CPU/test_loc_Segmentation2D.cpp:120-135).
    h, w = 3 * 130, 3 * 130
    w1 = np.zeros((h, w))
    w1[0::3, 0::3] = 1.0
    w1[1::3, 0::3] = 1.0
    w1[0::3, 1::3] = 1.0
    w1[1::3, 1::3] = 2.0
    w1[2::3, 0::3] = 0.5   # fifth pixel: surface 5 > VOLUME_MIN
    _, reg = oracle_mod.segment(w1, 0.1)
    assert len(reg) == 130 * 130


@pytest.mark.parametrize("name", ["small", "dense"])
@pytest.mark.parametrize("ratio_idx", [0, 1])
def test_segment_labels_match_compiled_reference(oracle_mod, ref_vectors, name, ratio_idx):
    st = ref_vectors[name]
    thr = float(ref_vectors[f"{name}_autothreshold"][0])
    labels = ref_vectors[f"{name}_labels"][ratio_idx]
    regs = ref_vectors[f"{name}_regions"]
    for f in range(st.shape[0]):
        w1 = oracle_mod.atrous(st[f])[1]
        lab, reg = oracle_mod.segment(w1, thr, (0.0, 10.0)[ratio_idx])
        assert np.array_equal(lab, labels[f])
        exp = regs[(regs[:, 0] == ratio_idx) & (regs[:, 1] == f)][:, 2:]
        assert np.array_equal(reg[:, [0, 1, 3, 4]], exp)


def test_dense_frames_have_multi_label_components(oracle_mod, ref_vectors):
    """The dense fixture exercises touching spots (components holding several labels)."""
    from scipy import ndimage
    st = ref_vectors["dense"]
    thr = float(ref_vectors["dense_autothreshold"][0])
    multi = 0
    for f in range(st.shape[0]):
        w1 = oracle_mod.atrous(st[f])[1]
        lab = oracle_mod.watershed_labels(w1, thr)
        cc, n = ndimage.label(lab > 0, structure=np.ones((3, 3)))
        for c in range(1, n + 1):
            multi += len(np.unique(lab[cc == c])) > 1
    assert multi >= 10


# ---------------------------------------------------------------- linear algebra ----------------
def test_gauss_jordan_inverse(oracle_mod):
    rng = np.random.default_rng(5)
    for n in range(1, 8):
        a = rng.standard_normal((n, n)) + n * np.eye(n)
        ok, inv = oracle_mod.inverse_gauss_jordan(a)
        assert ok
        assert np.allclose(inv @ a, np.eye(n), atol=1e-12)
        assert np.allclose(inv, np.linalg.inv(a), rtol=1e-10, atol=1e-12)
    ok, inv = oracle_mod.inverse_gauss_jordan([[2.0, 0.0], [0.0, 4.0]])
    assert ok and np.array_equal(inv, [[0.5, 0.0], [0.0, 0.25]])
    # Pivoting: a zero leading entry (Tests_CPU/test_sp_Matrix.cpp:230-298).
    ok, inv = oracle_mod.inverse_gauss_jordan([[0.0, 1.0], [1.0, 0.0]])
    assert ok and np.array_equal(inv, [[0.0, 1.0], [1.0, 0.0]])
    for bad in ([[1.0, 2.0], [2.0, 4.0]], [[0.0, 0.0], [1.0, 1.0]], [[np.nan, 1.0], [0.0, 1.0]]):
        ok, inv = oracle_mod.inverse_gauss_jordan(bad)
        assert not ok and not inv.any()


# ---------------------------------------------------------------- Gaussian fit ------------------
def _point_gaussian(h, w, x0, y0, a, bg, sx, sy=None, theta=0.0):
    """Point-sampled Gaussian rounded to u16 (Tests_CPU/test_fit_Fitting.cpp:36-66)."""
    sy = sx if sy is None else sy
    yy, xx = np.mgrid[0:h, 0:w].astype(np.float64)
    dx, dy = xx - x0, yy - y0
    c, s = np.cos(theta), np.sin(theta)
    f1, f2 = (dx * c - dy * s) / sx, (dx * s + dy * c) / sy
    return np.round(bg + a * np.exp(-(f1 * f1 + f2 * f2) / 2)).astype(np.uint16)


@pytest.mark.parametrize("mode", [1, 2, 3, 4])
def test_fit_recovers_point_gaussian(oracle_mod, mode):
    img = _point_gaussian(14, 15, 7.25, 6.7, 8000.0, 100.0, 1.5)
    row = oracle_mod.fit_region(mode, img, (7, 1.5, 1.5, 0.0), 7.0, 7.0, seg=1.0, surf=9.0)
    assert abs(row[X] - 7.25) < 0.01 and abs(row[Y] - 6.7) < 0.01     # test_fit_Fitting.cpp:236-256
    if mode >= 2:
        assert abs(row[SX] - 1.5) < 0.02 and abs(row[SY] - 1.5) < 0.02
    assert abs(row[13] - 8000) < 5 and abs(row[14] - 100) < 2
    assert row[QXY] >= 0.99999


def test_fit_r2_is_squared_pearson(oracle_mod):
    rng = np.random.default_rng(9)
    img = _point_gaussian(20, 20, 9.6, 10.2, 600.0, 300.0, 1.3) + rng.integers(0, 40, (20, 20)).astype(np.uint16)
    row = oracle_mod.fit_region(2, img, (7, 1.3, 1.3, 0.0), 9.6, 10.2)
    # Recompute r^2 from the fitted parameters and the 7x7 patch (test_fit_Fitting.cpp:259-288).
    ux, uy = int(9.6), int(10.2)
    yy, xx = np.mgrid[uy - 3:uy + 4, ux - 3:ux + 4].astype(np.float64)
    s = row[SX]
    model = row[14] + row[13] * np.exp(-((xx - row[X]) ** 2 + (yy - row[Y]) ** 2) / (2 * s * s))
    data = img[uy - 3:uy + 4, ux - 3:ux + 4].astype(np.float64)
    r = np.corrcoef(data.ravel(), model.ravel())[0, 1]
    assert row[QXY] == pytest.approx(r * r, abs=1e-10)


def test_fit_sentinels(oracle_mod):
    img = _point_gaussian(20, 20, 10.0, 10.0, 800.0, 100.0, 1.3)
    border = oracle_mod.fit_region(2, img, (7, 1.3, 1.3, 0), 2.5, 10.0)
    assert border[X] == 2.5 and border[13] == -1 and border[14] == -1 and border[INTENSITY] == -1 and border[CIRC] == 1
    nan = oracle_mod.fit_region(2, img, (7, np.nan, 1.3, 0), 10.0, 10.0)
    assert nan[13] == -1 and nan[SX] == -1
    flat = oracle_mod.fit_region(2, np.full((20, 20), 500, np.uint16), (7, 1.3, 1.3, 0), 10.0, 10.0)
    assert flat[QXY] == -1 or flat[13] == -1


# ---------------------------------------------------------------- end to end --------------------
@pytest.mark.parametrize("fit", [0, 1, 2, 3, 4])
@pytest.mark.parametrize("watershed", [True, False])
def test_localization_matches_golden_csv(oracle_mod, ref_inputs, ref_csv, fit, watershed):
    st = ref_inputs["stack"]
    rows = oracle_mod.localization(st, 103.6, 0.0 if watershed else 10.0, fit, FIT_PARAMS_CSV)
    exp = ref_csv[f"stack-localizations-103.6_{watershed}_{fit}_1.0_0.0_7"]
    assert match_rows(theta_to_degrees(rows, fit), exp, atol=1e-3) == []


def test_quadrant_matches_golden_csv(oracle_mod, ref_inputs, ref_csv):
    rows = oracle_mod.localization(ref_inputs["stack_quadrant"], 103.6, 0.0, 4, FIT_PARAMS_CSV)
    exp = ref_csv["stack_quadrant-localizations-103.6_True_4_1.0_0.0_7"]
    assert match_rows(theta_to_degrees(rows, 4), exp, atol=1e-5) == []


def test_autothreshold_known_answers(oracle_mod, ref_inputs):
    expected = [63.651749, 66.142910, 63.192296, 64.368855, 63.954150,
                63.400043, 66.521994, 63.373237, 62.613244, 63.866017]  # Tests_CPU/test_Palm.cpp:307-310
    st = ref_inputs["stack"]
    for f in range(10):
        assert oracle_mod.autothreshold(st[f], (7.0,)) == pytest.approx(expected[f], abs=1e-5)


@pytest.mark.parametrize("name", ["small", "dense"])
def test_localization_bit_exact_vs_compiled_reference(oracle_mod, ref_vectors, name):
    st = ref_vectors[name]
    thr = float(ref_vectors[f"{name}_autothreshold"][0])
    assert oracle_mod.autothreshold(st[0]) == thr
    for fit in range(5):
        for ratio in (0, 10):
            rows = oracle_mod.localization(st, thr, float(ratio), fit, (7, 1.3, 1.3, 0))
            exp = ref_vectors[f"{name}_loc_fit{fit}_ratio{ratio}"]
            assert rows.shape == exp.shape
            assert np.array_equal(rows, exp), (fit, ratio, np.nanmax(np.abs(rows - exp)))
    pos = ref_vectors[f"{name}_positions"]
    for fit in range(1, 5):
        assert np.array_equal(oracle_mod.fitting(st, pos, fit, (7, 1.3, 1.3, 0)), ref_vectors[f"{name}_fitting_fit{fit}"])


def test_fitting_matches_localization(oracle_mod, ref_inputs):
    # Tests_CPU/test_Palm.cpp:269-295.
    st = ref_inputs["stack"]
    det = oracle_mod.localization(st, 103.6, 0.0, 0, FIT_PARAMS_CSV)
    exp = oracle_mod.localization(st, 103.6, 0.0, 2, FIT_PARAMS_CSV)
    pos = np.column_stack([2 * det[:, ID], det[:, PLANE], det[:, X], det[:, Y], det[:, 15], det[:, 16]])
    out = oracle_mod.fitting(st, pos, 2, FIT_PARAMS_CSV)
    assert np.array_equal(out[:, ID], pos[:, 0])
    a, e = out.copy(), exp.copy()
    a[:, [ID, INDEX]] = 0
    e[:, [ID, INDEX]] = 0
    assert match_rows(a, e, atol=1e-8) == []


def test_per_plane_capacity(oracle_mod, ref_inputs):
    # Tests_CPU/test_Palm.cpp:105-154: capacity (out/18)/planes per plane, kept in label order.
    st = ref_inputs["stack"]
    z = st.shape[0]
    rows = oracle_mod.localization(st, 103.6, 0.0, 0, FIT_PARAMS_CSV, output_size=2 * 18 * z + 1)
    assert rows.shape[0] == 2 * z
    assert np.array_equal(rows[:, PLANE], np.repeat(np.arange(1, z + 1), 2))
    assert np.array_equal(rows[:, INDEX], np.tile([1, 2], z))
    w1 = oracle_mod.atrous(st[0])[1]
    _, reg = oracle_mod.segment(w1, 103.6)
    first_two = reg[:2]
    got = rows[rows[:, PLANE] == 1][:, [X, Y]]
    assert sorted(map(tuple, got)) == sorted(map(tuple, first_two[:, :2]))


def test_empty_and_black(oracle_mod):
    assert oracle_mod.localization(np.zeros((3, 12, 16), np.uint16), 0.0, 0.0, 0).shape == (0, 18)
    assert oracle_mod.localization(np.zeros((2, 16, 16), np.uint16), 0.0, 0.0, 0, output_size=0).shape == (0, 18)


# ---------------------------------------------------------------- astigmatic z ------------------
# Models and cases of Tests_CPU/test_fit_AstigFit.cpp:20-54 (pixel size 160 nm, sampling 1).
AST_PIXEL = 160.0
AST_REF_MODEL = [-60.0, 300.0, 0.06, 0.02, 140.0, 60.0, 300.0, -0.06, 0.02, 140.0]
AST_EST_MODEL = [-1000.0, 1000.0, 0.0, 0.0, 160.0, 1000.0, 1000.0, 0.0, 0.0, 160.0]
AST_CASES = [(-600.0, -500.0), (-500.0, -500.0), (-499.0, -499.0), (-123.4, -123.4), (0.0, 0.0),
             (123.4, 123.4), (499.0, 499.0), (500.0, 500.0), (600.0, 500.0)]


def _ast_sigma(c, z, pixel=AST_PIXEL):
    """SigmaModel of Tests_CPU/test_fit_AstigFit.cpp:93-101, numpy."""
    u = (np.asarray(z, np.float64) - c[0]) / c[1]
    poly = 1.0 + u * u + c[2] * u ** 3 + c[3] * u ** 4
    return (1.0 / pixel) * c[4] * np.sqrt(np.maximum(0.0, poly))


def _ast_pairs(model, z):
    return np.stack([_ast_sigma(model[:5], z), _ast_sigma(model[5:], z)], axis=-1)


def test_astig_estimate_axial_limits(oracle_mod):
    """EstimateWithinLimits (test_fit_AstigFit.cpp:241-272): bounds are exact, interior points within
    1 nm, the score is the closed form at the returned z, and Brent improves on the 5 nm grid."""
    obs = np.array([c[0] for c in AST_CASES])
    sig = _ast_pairs(AST_EST_MODEL, obs)
    res = oracle_mod.astig_estimate(sig, AST_PIXEL, AST_EST_MODEL, 500.0)
    for (o, e), (z, q), s in zip(AST_CASES, res, sig):
        if abs(e) == 500.0:
            assert z == e
        else:
            assert abs(z - e) <= 1.0 and -500.0 < z < 500.0
        p = _ast_pairs(AST_EST_MODEL, z)
        r = (p - s) / p
        assert abs(q - 1.0 / (1.0 + (r[0] ** 2 + r[1] ** 2) / (2 * 0.2 * 0.2))) <= 1e-12
        assert 0.0 <= q <= 1.0
        if abs(o) > 500.0:
            assert q < 1.0
        if abs(o) == 123.4:
            g = np.round(o / 5.0) * 5.0
            dg = _ast_pairs(AST_EST_MODEL, g) - s
            assert ((p - s) ** 2).sum() < (dg ** 2).sum()


@pytest.mark.parametrize("rel,score", [(0.0, 1.0), (0.01, 400.0 / 401.0), (0.05, 16.0 / 17.0), (0.10, 0.8),
                                       (0.20, 0.5), (1.0, 1.0 / 26.0), (-0.20, 0.5)])
def test_astig_quality_scale(oracle_mod, rel, score):
    """RelativeSigmaScore (test_fit_AstigFit.cpp:299-319): both observed sigmas off by the same
    relative error at z = 123.4."""
    z = 123.4
    obs = _ast_pairs(AST_EST_MODEL, z) * (1 + rel)
    assert abs(oracle_mod.astig_quality(AST_EST_MODEL, z, obs[0], obs[1], AST_PIXEL) - score) <= 1e-14


def test_astig_scale_invariance_and_invalid_widths(oracle_mod):
    """QualityScaleInvariant and InvalidWidthQuality (test_fit_AstigFit.cpp:322-346)."""
    s = _ast_pairs(AST_EST_MODEL, 123.4) * np.array([1.05, 0.98])
    a = oracle_mod.astig_estimate(s[None], AST_PIXEL, AST_EST_MODEL, 500.0)[0]
    b = oracle_mod.astig_estimate(2 * s[None], AST_PIXEL / 2, AST_EST_MODEL, 500.0)[0]
    assert a[0] == b[0] and a[1] == b[1]
    r = s / _ast_pairs(AST_EST_MODEL, a[0]) - 1
    assert abs(a[1] - 1.0 / (1.0 + (r ** 2).sum() / 0.08)) <= 1e-14
    for sx, sy in ((0.0, 1.0), (1.0, -1.0), (np.nan, 1.0)):
        assert oracle_mod.astig_quality(AST_EST_MODEL, 0.0, sx, sy, AST_PIXEL) == 0.0
    bad = np.array([[0.0, 1.0], [1.0, -1.0], [np.nan, 1.0]])
    assert np.all(oracle_mod.astig_estimate(bad, AST_PIXEL, AST_EST_MODEL, 500.0)[:, 1] == 0.0)
    inv = list(AST_EST_MODEL)
    inv[4] = 0.0
    assert oracle_mod.astig_quality(inv, 0.0, 1.0, 1.0, AST_PIXEL) == 0.0


def test_astig_estimate_recovers_z(oracle_mod):
    """Estimate (test_fit_AstigFit.cpp:223-238): noisy pairs from the calibrated model within 40 nm,
    score above 0.95."""
    rng = np.random.default_rng(42)
    z = np.linspace(-200.0, 200.0, 10)
    sig = _ast_pairs(AST_REF_MODEL, z) + rng.normal(0.0, 0.02, (10, 2))
    res = oracle_mod.astig_estimate(sig, AST_PIXEL, AST_REF_MODEL, 800.0)
    assert np.all(np.abs(res[:, 0] - z) <= 40.0)
    assert np.all((res[:, 1] > 0.95) & (res[:, 1] <= 1.0))


def test_astig_estimate_bit_exact_vs_compiled_reference(oracle_mod):
    rng = np.random.default_rng(11)
    n = 4000
    z = rng.uniform(-700.0, 700.0, n)
    sig = _ast_pairs(AST_REF_MODEL, z) * rng.normal(1.0, 0.05, (n, 2))
    sig[:50] = rng.normal(0.0, 1.0, (50, 2))           # EstimateBad: random pairs
    sig[50:53] = [[0.0, 1.0], [np.nan, 1.0], [np.inf, 2.0]]
    for model, zmax in ((AST_REF_MODEL, 600.0), (AST_EST_MODEL, 500.0), (AST_REF_MODEL, 50.0)):
        exp = oracle_mod.ref_astig_estimate(sig, AST_PIXEL, model, zmax)
        if exp is None:
            pytest.skip("compiled reference not built (oracle/_ref)")
        got = oracle_mod.astig_estimate(sig, AST_PIXEL, model, zmax)
        assert np.array_equal(got, exp, equal_nan=True)


# ---------------------------------------------------------------- spline fit (mode 5) -----------
def test_spline_fit_bit_exact_vs_compiled_reference(oracle_mod):
    """SplineFit (lib/fit/SplineFit.cpp) restated: Localization and Fitting with fit mode 5 equal the
    compiled reference's rows bit for bit (NaN where the reference produces NaN)."""
    import synth_np
    from helpers import spline_params
    st = synth_np.frames(3, 96, 128, 14, seed=21)
    fp = spline_params()
    thr = oracle_mod.autothreshold(st[0])
    exp = oracle_mod.ref_localization(st, thr, 0.0, 5, fp)
    if exp is None:
        pytest.skip("compiled reference not built (oracle/_ref)")
    got = oracle_mod.localization(st, thr, 0.0, 5, fp)
    assert got.shape == exp.shape and got.shape[0] > 20
    assert np.array_equal(got, exp, equal_nan=True)
    pos = np.stack([got[:, ID], got[:, PLANE], got[:, X] + 0.3, got[:, Y] - 0.2, got[:, 15], got[:, 16]], axis=1)
    exp2 = oracle_mod.ref_fitting(st, pos, 5, fp)
    got2 = oracle_mod.fitting(st, pos, 5, fp)
    assert np.array_equal(got2, exp2, equal_nan=True)


@pytest.mark.parametrize("size,roi", [((11, 11, 21), 7), ((13, 13, 9), 9), ((9, 9, 5), 7), ((11, 11, 21), 8),
                                      ((7, 7, 4), 9)])
def test_spline_fit_shapes_and_crlb(oracle_mod, size, roi):
    """Other spline and ROI sizes (the offset between them, the clamped spline indices) against the
    compiled reference; the CRLB columns are square roots (>= 0) where finite."""
    import synth_np
    from helpers import spline_params
    st = synth_np.frames(2, 64, 64, 8, seed=33)
    fp = spline_params(roi=roi, size=size, seed=5)
    thr = oracle_mod.autothreshold(st[0])
    got = oracle_mod.localization(st, thr, 0.0, 5, fp)
    exp = oracle_mod.ref_localization(st, thr, 0.0, 5, fp)
    if exp is not None:
        assert np.array_equal(got, exp, equal_nan=True)
    fitted = got[got[:, SX] == 0]                   # rows of the fit (border rows carry -1)
    for c in (11, 13, 14):                         # CRLB columns: QualityXY, Amplitude, Offset
        v = fitted[:, c]
        assert np.all(v[np.isfinite(v)] >= 0)
