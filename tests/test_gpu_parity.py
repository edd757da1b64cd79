"""Parity of the HIP path (through the C ABI) with the CPU oracle and the reference's golden data."""
import numpy as np
import pytest

from helpers import (AMP, ID, INDEX, PLANE, SEG, SURF, X, Y, fit_errors, filter_small, fit_mismatches, match_rows,
                     rank_labels, theta_to_degrees)

pytestmark = pytest.mark.gpu
FIT_PARAMS_CSV = (7.0, 1.0, 2.0, 0.0)
FIT_PARAMS_SYN = (7.0, 1.3, 1.3, 0.0)


def assert_fp64_fit(got, exp, fit):
    """FP64 fits run the reference's arithmetic in its order with exp, sin and cos bit-equal to the C
    library's, so every row equals the oracle's bit for bit."""
    assert got.shape == exp.shape
    same = np.all((got == exp) | (np.isnan(got) & np.isnan(exp)), axis=1)
    assert same.all(), (fit, np.flatnonzero(~same)[:5], fit_errors(got, exp))


@pytest.fixture(autouse=True)
def fp64_default(pt_gpu):
    pt_gpu.set_option("fit_fp64", 1)
    pt_gpu.set_option("max_iter", 75)
    yield


def _stacks(ref_inputs, ref_vectors):
    return {
        "stack": (ref_inputs["stack"], 103.6),
        "small": (ref_vectors["small"], float(ref_vectors["small_autothreshold"][0])),
        "dense": (ref_vectors["dense"], float(ref_vectors["dense_autothreshold"][0])),
    }


# ------------------------------------------------------------------ a trous -----------------------
@pytest.mark.parametrize("level", [0, 1, 2, 3, 4])
def test_wavelett_bit_exact(pt_gpu, oracle_mod, ref_inputs, level):
    st = ref_inputs["stack"]
    got = pt_gpu.wavelett(st, level)
    exp = oracle_mod.wavelett(st, level)
    assert np.array_equal(got, exp)


@pytest.mark.parametrize("shape", [(1, 1), (9, 1), (1, 9), (13, 7), (65, 33), (3, 700), (300, 2)])
def test_wavelett_shapes(pt_gpu, oracle_mod, shape):
    rng = np.random.default_rng(sum(shape))
    st = rng.integers(0, 65536, (2,) + shape, dtype=np.uint16)
    for level in (1, 2, 3):
        assert np.array_equal(pt_gpu.wavelett(st, level), oracle_mod.wavelett(st, level))


def test_wavelett_saturated_extremes(pt_gpu, oracle_mod):
    st = np.zeros((1, 40, 40), np.uint16)
    st[0, ::2, ::3] = 65535
    for level in (1, 2, 3):
        assert np.array_equal(pt_gpu.wavelett(st, level), oracle_mod.wavelett(st, level))


# ------------------------------------------------------------------ watershed labels --------------
@pytest.mark.parametrize("name", ["stack", "small", "dense"])
def test_labels_bit_exact(pt_gpu, oracle_mod, ref_inputs, ref_vectors, name):
    st, thr = _stacks(ref_inputs, ref_vectors)[name]
    for f in range(st.shape[0]):
        w1 = oracle_mod.atrous(st[f])[1]
        exp = oracle_mod.watershed_labels(w1, thr)
        got = pt_gpu.debug_labels(st[f], thr)
        assert np.array_equal(got > 0, exp > 0), f"mask differs in frame {f}"
        assert np.array_equal(rank_labels(got), exp), f"labels differ in frame {f}"


def test_labels_threshold_edges(pt_gpu, oracle_mod, ref_vectors):
    """Thresholds placed exactly on W1 values: the strict comparison must match bit for bit."""
    st = ref_vectors["dense"]
    w1 = oracle_mod.atrous(st[0])[1]
    vals = np.unique(w1[w1 > 5])
    for thr in (vals[len(vals) // 3], vals[len(vals) // 2], -1.0, 0.0):
        exp = oracle_mod.watershed_labels(w1, float(thr))
        got = pt_gpu.debug_labels(st[0], float(thr))
        assert np.array_equal(rank_labels(got), exp)


def _odd_frame(h, w, seed):
    """Spots of several widths, one broad blob (long mask runs, several maxima) and shot noise."""
    rng = np.random.default_rng(seed)
    yy, xx = np.mgrid[0:h, 0:w].astype(np.float64)
    img = np.full((h, w), 100.0)
    for _ in range(max(1, h * w // 150)):
        cy, cx = rng.uniform(-2, h + 1), rng.uniform(-2, w + 1)
        sg, amp = rng.uniform(1.0, 2.5), rng.uniform(300, 3000)
        img += amp * np.exp(-((yy - cy) ** 2 + (xx - cx) ** 2) / (2 * sg * sg))
    img += 1500 * np.exp(-((yy - h / 2) ** 2 + (xx - w / 3) ** 2) / (2 * 6.0 ** 2))
    return np.clip(rng.poisson(img), 0, 65535).astype(np.uint16)


@pytest.mark.parametrize("shape", [(1, 1), (1, 70), (70, 1), (23, 37), (33, 65), (61, 97), (7, 300), (90, 91)])
def test_labels_odd_shapes(pt_gpu, oracle_mod, shape):
    """Frame widths and pixel counts that are not multiples of the bitmap words: mask runs that start
    inside a word and cross word and row ends, at sparse, dense and full masks."""
    h, w = shape
    fr = _odd_frame(h, w, 1000 + h * 7 + w)
    w1 = oracle_mod.atrous(fr)[1]
    for thr in (float(np.percentile(w1, 90)), float(np.percentile(w1, 50)), -1e9):
        exp = oracle_mod.watershed_labels(w1, thr)
        got = pt_gpu.debug_labels(fr, thr)
        assert np.array_equal(got > 0, exp > 0), f"mask differs at {shape} thr {thr}"
        assert np.array_equal(rank_labels(got), exp), f"labels differ at {shape} thr {thr}"
    st = fr[None]
    thr = float(np.percentile(w1, 80))
    assert np.array_equal(pt_gpu.localization(st, thr, 0.0, 0, FIT_PARAMS_SYN),
                          oracle_mod.localization(st, thr, 0.0, 0, FIT_PARAMS_SYN))


# ------------------------------------------------------------------ localization ------------------
@pytest.mark.parametrize("name", ["stack", "small", "dense"])
def test_localization_no_fit_bit_exact(pt_gpu, oracle_mod, ref_inputs, ref_vectors, name):
    st, thr = _stacks(ref_inputs, ref_vectors)[name]
    got = pt_gpu.localization(st, thr, 0.0, 0, FIT_PARAMS_SYN)
    exp = oracle_mod.localization(st, thr, 0.0, 0, FIT_PARAMS_SYN)
    assert got.shape == exp.shape
    assert np.array_equal(got, exp)


@pytest.mark.parametrize("fit", [1, 2, 3, 4])
@pytest.mark.parametrize("name", ["small", "dense"])
def test_localization_fp64_matches_oracle(pt_gpu, oracle_mod, ref_vectors, name, fit):
    st, thr = ref_vectors[name], float(ref_vectors[f"{name}_autothreshold"][0])
    got = pt_gpu.localization(st, thr, 0.0, fit, FIT_PARAMS_SYN)
    exp = oracle_mod.localization(st, thr, 0.0, fit, FIT_PARAMS_SYN)
    assert_fp64_fit(got, exp, fit)


@pytest.mark.parametrize("fit", [1, 2, 3, 4])
@pytest.mark.parametrize("name", ["stack", "dense"])
def test_localization_fp64_bit_equal_to_compiled_reference(pt_gpu, oracle_mod, ref_inputs, ref_vectors, name, fit):
    """FP64 rows against the reference's own code (lib/*.cpp compiled in place by oracle/Makefile
    into oracle/_ref, which travels with the built library): bit for bit, every column."""
    if oracle_mod.ref_lib() is None:
        pytest.skip("the compiled reference (oracle/_ref) was not built")
    if name == "stack":
        st, thr, fp = ref_inputs["stack"], 103.6, FIT_PARAMS_CSV
    else:
        st, thr, fp = ref_vectors[name], float(ref_vectors[f"{name}_autothreshold"][0]), FIT_PARAMS_SYN
    got = pt_gpu.localization(st, thr, 0.0, fit, fp)
    exp = oracle_mod.ref_localization(st, thr, 0.0, fit, fp)
    assert got.shape == exp.shape and got.shape[0] > 100
    same = np.all((got == exp) | (np.isnan(got) & np.isnan(exp)), axis=1)
    assert same.all(), (np.flatnonzero(~same)[:5], fit_errors(got, exp))


@pytest.mark.parametrize("fit", [0, 1, 2, 3, 4])
def test_localization_matches_golden_csv(pt_gpu, ref_inputs, ref_csv, fit):
    rows = pt_gpu.localization(ref_inputs["stack"], 103.6, 0.0, fit, FIT_PARAMS_CSV)
    exp = ref_csv[f"stack-localizations-103.6_True_{fit}_1.0_0.0_7"]
    assert match_rows(theta_to_degrees(rows, fit), exp, atol=1e-3) == []   # Tests_CPU/test_Palm.cpp:55-76


def test_quadrant_matches_golden_csv(pt_gpu, ref_inputs, ref_csv):
    rows = pt_gpu.localization(ref_inputs["stack_quadrant"], 103.6, 0.0, 4, FIT_PARAMS_CSV)
    exp = ref_csv["stack_quadrant-localizations-103.6_True_4_1.0_0.0_7"]
    assert match_rows(theta_to_degrees(rows, 4), exp, atol=1e-5) == []


@pytest.mark.parametrize("fit", [1, 2, 3, 4])
def test_localization_fast_fit_within_tolerance(pt_gpu, oracle_mod, ref_vectors, fit):
    """Fast fit (FP32 kernel, FP64 hand-over; mode 4 in FP64): every row within the fit tolerance of
    the FP64 oracle (positions 1e-3 px, amplitude / offset / sigma 1e-4 relative)."""
    st, thr = ref_vectors["small"], float(ref_vectors["small_autothreshold"][0])
    pt_gpu.set_option("fit_fp64", 0)
    got = pt_gpu.localization(st, thr, 0.0, fit, FIT_PARAMS_SYN)
    exp = oracle_mod.localization(st, thr, 0.0, fit, FIT_PARAMS_SYN)
    assert got.shape == exp.shape
    ok = (exp[:, AMP] > 0) & (got[:, AMP] > 0)
    assert ok.mean() > 0.9
    bad = fit_mismatches(got, exp)
    assert bad == [], (bad[:5], fit_errors(got, exp))


def test_per_plane_capacity(pt_gpu, oracle_mod, ref_inputs):
    st = ref_inputs["stack"]
    z = st.shape[0]
    for cap in (2 * 18 * z + 1, 18 * z, 5 * 18 * z):
        got = pt_gpu.localization(st, 103.6, 0.0, 0, FIT_PARAMS_CSV, output_size=cap)
        exp = oracle_mod.localization(st, 103.6, 0.0, 0, FIT_PARAMS_CSV, output_size=cap)
        assert np.array_equal(got, exp)


def test_empty_inputs(pt_gpu):
    assert pt_gpu.localization(np.zeros((3, 12, 16), np.uint16), 0.0, 0.0, 0).shape == (0, 18)
    assert pt_gpu.localization(np.zeros((2, 16, 16), np.uint16), 0.0, 0.0, 2, output_size=0).shape == (0, 18)
    assert pt_gpu.fitting(np.zeros((1, 16, 16), np.uint16), np.zeros((0, 6)), 2).shape == (0, 18)


def test_output_capacity_exceeded(pt_gpu, frames512):
    """An output smaller than the rows of the call is an error (no rows returned) and nothing is
    written past it; with room for every row the device entry returns them all."""
    import torch
    st, thr = frames512
    d = torch.from_numpy(st).cuda()
    full = torch.zeros(d.shape[0] * 400 * 18, dtype=torch.float64, device="cuda")
    n = pt_gpu.localization_device(d, full, thr, 0.0, 2, FIT_PARAMS_SYN, 400)
    assert n > 10
    buf = torch.full(((n + 8) * 18,), -7.0, dtype=torch.float64, device="cuda")
    assert pt_gpu.localization_device(d, buf[: (n - 5) * 18], thr, 0.0, 2, FIT_PARAMS_SYN, 400) == 0
    assert bool((buf[(n - 5) * 18:] == -7.0).all())
    assert pt_gpu.localization_device(d, buf[: n * 18], thr, 0.0, 2, FIT_PARAMS_SYN, 400) == n
    assert torch.equal(buf[: n * 18], full[: n * 18]) and bool((buf[n * 18:] == -7.0).all())


def test_frames_are_independent(pt_gpu, ref_inputs):
    """Empty frames between useful ones change nothing (Tests_CPU/test_Palm.cpp:160-215)."""
    st = ref_inputs["stack"]
    single = pt_gpu.localization(st[:1], 103.6, 0.0, 4, FIT_PARAMS_CSV)
    z = np.zeros((5,) + st.shape[1:], np.uint16)
    z[1] = st[0]
    z[3] = st[0]
    got = pt_gpu.localization(z, 103.6, 0.0, 4, FIT_PARAMS_CSV)
    assert got.shape[0] == 2 * single.shape[0]
    n = single.shape[0]
    assert np.array_equal(got[:, ID], np.arange(1, 2 * n + 1))
    assert np.array_equal(got[:, PLANE], np.repeat([2, 4], n))
    assert np.array_equal(got[:n, 3:], single[:, 3:]) and np.array_equal(got[n:, 3:], single[:, 3:])


# ------------------------------------------------------------------ fit only ----------------------
@pytest.mark.parametrize("fit", [1, 2, 3, 4])
def test_fitting_matches_oracle(pt_gpu, oracle_mod, ref_vectors, fit):
    st = ref_vectors["dense"]
    pos = ref_vectors["dense_positions"]
    got = pt_gpu.fitting(st, pos, fit, FIT_PARAMS_SYN)
    exp = oracle_mod.fitting(st, pos, fit, FIT_PARAMS_SYN)
    assert np.array_equal(got[:, [ID, PLANE, INDEX]], exp[:, [ID, PLANE, INDEX]])
    assert_fp64_fit(got, exp, fit)


def test_fitting_matches_localization(pt_gpu, ref_inputs):
    st = ref_inputs["stack"]
    det = pt_gpu.localization(st, 103.6, 0.0, 0, FIT_PARAMS_CSV)
    exp = pt_gpu.localization(st, 103.6, 0.0, 2, FIT_PARAMS_CSV)
    pos = np.column_stack([2 * det[:, ID], det[:, PLANE], det[:, X], det[:, Y], det[:, SEG], det[:, SURF]])
    out = pt_gpu.fitting(st, pos, 2, FIT_PARAMS_CSV)
    a, e = out.copy(), exp.copy()
    a[:, [ID, INDEX]] = 0
    e[:, [ID, INDEX]] = 0
    assert match_rows(a, e, atol=1e-8) == []


# ------------------------------------------------------------------ device generator + full path --
def test_synthetic_stack_end_to_end(pt_gpu, oracle_mod):
    import torch
    d = torch.empty((4, 256, 256), dtype=torch.uint16, device="cuda")
    pt_gpu.synth_device(d, seed=21, emitters=30)
    st = d.cpu().numpy()
    assert 100 <= st.min() and st.mean() > 130
    thr = oracle_mod.autothreshold(st[0])
    got = pt_gpu.localization(st, thr, 0.0, 2, FIT_PARAMS_SYN)
    exp = oracle_mod.localization(st, thr, 0.0, 2, FIT_PARAMS_SYN)
    assert got.shape == exp.shape and got.shape[0] > 50
    assert_fp64_fit(got, exp, 2)
    # Device-resident entry: same rows.
    out = torch.empty(((256 * 256) // 5 + 1) * 4 * 18, dtype=torch.float64, device="cuda")
    n = pt_gpu.localization_device(d, out, thr, 0.0, 2, FIT_PARAMS_SYN)
    assert np.array_equal(out[: n * 18].view(-1, 18).cpu().numpy(), got)


# ------------------------------------------------------------------ basin fusion (ratio >= 1) ----
_filter_small = filter_small


@pytest.mark.parametrize("name", ["stack", "small", "dense"])
def test_fusion_labels_bit_exact(pt_gpu, oracle_mod, ref_inputs, ref_vectors, name):
    st, thr = _stacks(ref_inputs, ref_vectors)[name]
    for f in range(st.shape[0]):
        w1 = oracle_mod.atrous(st[f])[1]
        exp, _ = oracle_mod.segment(w1, thr, 10.0)
        got = pt_gpu.debug_labels(st[f], thr, 10.0)
        assert np.array_equal(_filter_small(got), exp), f"frame {f}"


@pytest.mark.parametrize("fit", [0, 1, 2, 3, 4])
def test_localization_without_watershed_matches_golden_csv(pt_gpu, ref_inputs, ref_csv, fit):
    rows = pt_gpu.localization(ref_inputs["stack"], 103.6, 10.0, fit, FIT_PARAMS_CSV)
    exp = ref_csv[f"stack-localizations-103.6_False_{fit}_1.0_0.0_7"]
    assert match_rows(theta_to_degrees(rows, fit), exp, atol=1e-3) == []   # Tests_CPU/test_Palm.cpp:55-76


@pytest.mark.parametrize("name", ["small", "dense"])
def test_localization_fusion_matches_oracle(pt_gpu, oracle_mod, ref_vectors, name):
    st, thr = ref_vectors[name], float(ref_vectors[f"{name}_autothreshold"][0])
    got = pt_gpu.localization(st, thr, 10.0, 0, FIT_PARAMS_SYN)
    exp = oracle_mod.localization(st, thr, 10.0, 0, FIT_PARAMS_SYN)
    assert np.array_equal(got, exp)
    got = pt_gpu.localization(st, thr, 10.0, 2, FIT_PARAMS_SYN)
    exp = oracle_mod.localization(st, thr, 10.0, 2, FIT_PARAMS_SYN)
    assert_fp64_fit(got, exp, 2)


@pytest.mark.parametrize("name", ["exp", "sin", "cos"])
def test_fit_libm_bit_equal_to_c_library(pt_gpu, oracle_mod, name):
    """The fit model's exp, sin and cos on the device return the oracle's C-library values bit for
    bit: over the fit's argument range, every reduction path and special values."""
    import torch
    rng = np.random.default_rng(0x5EED0E)
    if name == "exp":
        specials = [0.0, -0.0, 1e-300, -1e-300, 2.0 ** -54, -(2.0 ** -55), 512.0, -512.0, 709.782712893384, 709.79,
                    -745.1332191019411, -745.14, -1024.0, 1024.0, np.inf, -np.inf, np.nan]
        parts = [rng.uniform(-40.0, 1.0, 1 << 22), rng.uniform(-760.0, 720.0, 1 << 20), rng.uniform(-745.2, -700.0, 1 << 18)]
    else:
        specials = [0.0, -0.0, 2.0 ** -27, -(2.0 ** -26), 0.126, -0.126, 0.85546875, -0.8554687, 2.426265, -2.4262651,
                    np.pi / 2, np.pi, 105414349.0, -105414349.0, np.inf, -np.inf, np.nan]
        parts = [rng.uniform(-0.126, 0.126, 1 << 20), rng.uniform(-2.43, 2.43, 1 << 22), rng.uniform(-100.0, 100.0, 1 << 20),
                 rng.uniform(-1e8, 1e8, 1 << 18)]
    x = np.concatenate(parts + [np.array(specials)])
    d_x = torch.from_numpy(x).cuda()
    d_y = torch.empty_like(d_x)
    pt_gpu.debug_libm(name, d_x, d_y)
    got = d_y.cpu().numpy()
    exp = oracle_mod.libm(name, x)
    same = (got.view(np.uint64) == exp.view(np.uint64)) | (np.isnan(got) & np.isnan(exp))
    assert same.all(), list(zip(x[~same][:5], got[~same][:5], exp[~same][:5]))


# ------------------------------------------------------------------ AutoThreshold ------------------
def test_autothreshold_known_answers(pt_gpu, ref_inputs):
    expected = [63.651749, 66.142910, 63.192296, 64.368855, 63.954150,
                63.400043, 66.521994, 63.373237, 62.613244, 63.866017]  # Tests_CPU/test_Palm.cpp:307-310
    st = ref_inputs["stack"]
    for f in range(10):
        assert pt_gpu.autothreshold(st[f], (7.0,)) == pytest.approx(expected[f], abs=1e-5)


@pytest.mark.parametrize("name", ["small", "dense"])
def test_autothreshold_matches_compiled_reference(pt_gpu, ref_vectors, name):
    thr = float(ref_vectors[f"{name}_autothreshold"][0])
    assert pt_gpu.autothreshold(ref_vectors[name][0], (7.0,)) == pytest.approx(thr, rel=1e-12)


# ------------------------------------------------------------------ 512^2 frames, both flood paths --
@pytest.fixture(scope="module")
def frames512(pt_gpu, oracle_mod):
    import torch
    d = torch.empty((12, 512, 512), dtype=torch.uint16, device="cuda")
    pt_gpu.synth_device(d, seed=0x5EED0002, emitters=200)
    st = d.cpu().numpy()
    return st, oracle_mod.autothreshold(st[0])


@pytest.mark.parametrize("sequential", [0, 1, 2, 3, 4])
@pytest.mark.parametrize("ratio", [0.0, 10.0])
def test_labels_512(pt_gpu, oracle_mod, frames512, sequential, ratio):
    """0: component floods; every frame through the exact frame flood: 1 grid rank flood (then the
    single-workgroup floods for frames it hands over), 2 wave-queue flood, 3 heap restatement,
    4 single-workgroup rank flood."""
    st, thr = frames512
    pt_gpu.set_option("force_sequential_watershed", sequential > 0)
    pt_gpu.set_option("force_wave_flood", sequential == 2)
    pt_gpu.set_option("force_heap_watershed", sequential == 3)
    pt_gpu.set_option("force_rank_flood", sequential == 4)
    try:
        for f in range(st.shape[0]):
            w1 = oracle_mod.atrous(st[f])[1]
            got = pt_gpu.debug_labels(st[f], thr, ratio)
            assert pt_gpu.last_stats()["status"] == 0
            if ratio == 0.0:
                exp = oracle_mod.watershed_labels(w1, thr)
                ok = np.array_equal(rank_labels(got), exp)
            else:
                exp, _ = oracle_mod.segment(w1, thr, ratio)
                ok = np.array_equal(_filter_small(got), exp)
            assert ok, f"frame {f}: {int((got > 0).sum())} labelled pixels, first difference at " \
                       f"{np.argwhere((rank_labels(got) if ratio == 0.0 else _filter_small(got)) != exp)[:1].tolist()}"
    finally:
        pt_gpu.set_option("force_sequential_watershed", 0)
        pt_gpu.set_option("force_heap_watershed", 0)
        pt_gpu.set_option("force_wave_flood", 0)
        pt_gpu.set_option("force_rank_flood", 0)


def test_localization_512_no_fit(pt_gpu, oracle_mod, frames512):
    st, thr = frames512
    got = pt_gpu.localization(st, thr, 0.0, 0, FIT_PARAMS_SYN)
    exp = oracle_mod.localization(st, thr, 0.0, 0, FIT_PARAMS_SYN)
    assert got.shape == exp.shape and np.array_equal(got, exp)


@pytest.mark.parametrize("ratio", [0.0, 10.0])
def test_dense_and_sparse_frames_in_one_chunk(pt_gpu, oracle_mod, frames512, ratio):
    """Frames whose masks exceed the per-frame LDS capacity (about 8K pixels at 512^2) take the
    global-memory component labelling; sparse frames of the same launch stay in LDS."""
    st, thr = frames512
    rng = np.random.default_rng(7)
    dense = np.clip(st[1:3].astype(np.int64) + rng.integers(0, 3000, size=st[1:3].shape), 0, 65535).astype(np.uint16)
    stack = np.ascontiguousarray(np.stack([st[0], dense[0], st[3], dense[1]]))
    sizes = [int((oracle_mod.atrous(stack[f])[1] > thr).sum()) for f in range(stack.shape[0])]
    assert sizes[1] > 20000 and sizes[3] > 20000 and sizes[0] < 8000 and sizes[2] < 8000, sizes
    got = pt_gpu.localization(stack, thr, ratio, 0, FIT_PARAMS_SYN)
    exp = oracle_mod.localization(stack, thr, ratio, 0, FIT_PARAMS_SYN)
    assert got.shape == exp.shape and np.array_equal(got, exp)
    for f in (1, 3):
        w1 = oracle_mod.atrous(stack[f])[1]
        lab = pt_gpu.debug_labels(stack[f], thr, ratio)
        if ratio == 0.0:
            assert np.array_equal(rank_labels(lab), oracle_mod.watershed_labels(w1, thr)), f
        else:
            assert np.array_equal(_filter_small(lab), oracle_mod.segment(w1, thr, ratio)[0]), f


def test_partial_last_chunk(pt_gpu, oracle_mod, frames512):
    """A last chunk with fewer frames than the chunk capacity starts from clean per-frame state."""
    st, thr = frames512
    exp = oracle_mod.localization(st, thr, 0.0, 0, FIT_PARAMS_SYN)
    try:
        for frames_per_chunk in (5, 7, 11):
            pt_gpu.set_option("chunk_pixels", frames_per_chunk * 512 * 512)
            got = pt_gpu.localization(st, thr, 0.0, 0, FIT_PARAMS_SYN)
            assert got.shape == exp.shape and np.array_equal(got, exp), frames_per_chunk
    finally:
        pt_gpu.set_option("chunk_pixels", 0)


@pytest.mark.parametrize("fit", [1, 2, 3, 4])
def test_fast_fit_512(pt_gpu, oracle_mod, frames512, fit):
    """Fast fit over thousands of ROIs (lane refills, every LM exit, the hand-over to the FP64 group
    kernel; mode 4 whole in that kernel) against the FP64 oracle: every row within the tolerance."""
    st, thr = frames512
    pt_gpu.set_option("fit_fp64", 0)
    got = pt_gpu.localization(st, thr, 0.0, fit, FIT_PARAMS_SYN)
    exp = oracle_mod.localization(st, thr, 0.0, fit, FIT_PARAMS_SYN)
    assert got.shape == exp.shape and len(exp) > 1000
    bad = fit_mismatches(got, exp)
    assert bad == [], (len({b[0] for b in bad}), bad[:5], fit_errors(got, exp))


@pytest.mark.parametrize("waves", [1, 3, 64])
def test_fp32_fit_queue_widths(pt_gpu, frames512, waves):
    """Rows do not depend on how many persistent waves share the ROI queue."""
    st, thr = frames512
    pt_gpu.set_option("fit_fp64", 0)
    base = pt_gpu.localization(st[:3], thr, 0.0, 2, FIT_PARAMS_SYN)
    try:
        pt_gpu.set_option("fit_waves", waves)
        got = pt_gpu.localization(st[:3], thr, 0.0, 2, FIT_PARAMS_SYN)
    finally:
        pt_gpu.set_option("fit_waves", 0)
    assert np.array_equal(got, base)


@pytest.mark.parametrize("name", ["stack", "small", "dense"])
@pytest.mark.parametrize("exact", [1, 2, 3, 4])
def test_exact_floods_bit_exact(pt_gpu, oracle_mod, ref_inputs, ref_vectors, name, exact):
    """Every frame through each exact frame flood (grid rank flood, wave queue, heap, single-workgroup
    rank flood): labels equal the oracle's."""
    st, thr = _stacks(ref_inputs, ref_vectors)[name]
    pt_gpu.set_option("force_sequential_watershed", 1)
    pt_gpu.set_option("force_wave_flood", exact == 2)
    pt_gpu.set_option("force_heap_watershed", exact == 3)
    pt_gpu.set_option("force_rank_flood", exact == 4)
    try:
        for f in range(st.shape[0]):
            w1 = oracle_mod.atrous(st[f])[1]
            exp = oracle_mod.watershed_labels(w1, thr)
            got = pt_gpu.debug_labels(st[f], thr, 0.0)
            assert np.array_equal(rank_labels(got), exp), f"frame {f}"
    finally:
        pt_gpu.set_option("force_sequential_watershed", 0)
        pt_gpu.set_option("force_heap_watershed", 0)
        pt_gpu.set_option("force_wave_flood", 0)
        pt_gpu.set_option("force_rank_flood", 0)


@pytest.mark.parametrize("defer", [1, 0, 2])
def test_localization_with_all_frames_exact(pt_gpu, oracle_mod, frames512, defer):
    """Localization with every frame through the exact frame flood (chunks of 5 frames): the flagged
    frames of the first chunks deferred to the side stream (defer 1), resolved in line (defer 0), or
    deferred until the counters of a chunk show more than 2 % flagged frames (defer 2: chunks 0 and 1
    are queued before any counters are read); those of the last chunk resolved in line; rows
    unchanged."""
    st, thr = frames512
    exp = oracle_mod.localization(st, thr, 0.0, 0, FIT_PARAMS_SYN)
    pt_gpu.set_option("force_sequential_watershed", 1)
    pt_gpu.set_option("chunk_pixels", 5 * 512 * 512)
    pt_gpu.set_option("defer_flagged", defer)
    try:
        got = pt_gpu.localization(st, thr, 0.0, 0, FIT_PARAMS_SYN)
        stats = pt_gpu.last_stats()
    finally:
        pt_gpu.set_option("force_sequential_watershed", 0)
        pt_gpu.set_option("chunk_pixels", 0)
        pt_gpu.set_option("defer_flagged", 2)
    last = st.shape[0] - 5 * ((st.shape[0] - 1) // 5)
    assert stats["flagged_frames"] == st.shape[0]
    assert stats["deferred_frames"] == {1: st.shape[0] - last, 0: 0, 2: 10}[defer]
    assert got.shape == exp.shape and np.array_equal(got, exp)


# Seeded 256x256 frames (tests/golden/synth_np.py, 60 emitters) whose flood has equal keys (siblings
# with one potential): one pair or several. The rank flood settles a frame whose pairs are independent
# (each pair's other order changes nothing but its two ranks); other frames go on to the wave-queue
# flood, then the heap restatement. Found by solving the flood on the CPU.
TIE_FRAMES = [(101, 5.0, 1), (102, 4.0, 1), (104, 4.0, 1), (101, 4.0, 2), (103, 4.0, 3)]


@pytest.mark.parametrize("grid", [1, 0])
@pytest.mark.parametrize("seed,thr,pairs", TIE_FRAMES)
def test_equal_keys_in_the_flood(pt_gpu, oracle_mod, seed, thr, pairs, grid):
    """The grid rank flood settles a frame whose equal keys cannot change a label and hands the
    others on; the single-workgroup rank flood (grid 0) solves independent pairs in both orders."""
    import synth_np
    frame = synth_np.frames(1, 256, 256, 60, seed=seed)[0]
    w1 = oracle_mod.atrous(frame)[1]
    exp = oracle_mod.watershed_labels(w1, thr)
    pt_gpu.set_option("force_sequential_watershed", 1)
    pt_gpu.set_option("force_rank_flood", 1 - grid)
    try:
        got = pt_gpu.debug_labels(frame, thr, 0.0)
        stats = pt_gpu.last_stats()
    finally:
        pt_gpu.set_option("force_sequential_watershed", 0)
        pt_gpu.set_option("force_rank_flood", 0)
    assert np.array_equal(rank_labels(got), exp)
    if grid and stats["grid_frames"] == 1:
        assert stats["grid_equal_keys"] == 1, stats
    elif pairs == 1:
        assert stats["rank_two_orders"] == 1, stats
    else:
        assert stats["rank_two_orders"] == 1 or (stats["rank_equal_keys"] == 1 and stats["heap_frames"] == 1), stats


# ------------------------------------------------------------------ astigmatic z estimation --------
AST_PIXEL = 160.0
AST_REF_MODEL = [-60.0, 300.0, 0.06, 0.02, 140.0, 60.0, 300.0, -0.06, 0.02, 140.0]
AST_EST_MODEL = [-1000.0, 1000.0, 0.0, 0.0, 160.0, 1000.0, 1000.0, 0.0, 0.0, 160.0]


def _ast_pairs(model, z):
    m = np.asarray(model, np.float64)
    out = []
    for c in (m[:5], m[5:]):
        u = (np.asarray(z, np.float64) - c[0]) / c[1]
        poly = 1.0 + u * u + c[2] * u ** 3 + c[3] * u ** 4
        out.append((1.0 / AST_PIXEL) * c[4] * np.sqrt(np.maximum(0.0, poly)))
    return np.stack(out, axis=-1)


def _ast_inputs(n, seed):
    rng = np.random.default_rng(seed)
    z = rng.uniform(-700.0, 700.0, n)
    sig = _ast_pairs(AST_REF_MODEL, z) * rng.normal(1.0, 0.05, (n, 2))
    sig[:50] = rng.normal(0.0, 1.0, (50, 2))                      # random pairs (EstimateBad)
    sig[50:54] = [[0.0, 1.0], [np.nan, 1.0], [np.inf, 2.0], [1.0, -1.0]]
    cases = [-600.0, -500.0, -499.0, -123.4, 0.0, 123.4, 499.0, 500.0, 600.0]
    sig[54:63] = _ast_pairs(AST_EST_MODEL, cases)                 # the reference's axial cases
    return sig


@pytest.mark.parametrize("model,zmax", [(AST_REF_MODEL, 600.0), (AST_EST_MODEL, 500.0), (AST_REF_MODEL, 50.0)])
def test_astig_estimate_bit_exact(pt_gpu, oracle_mod, model, zmax):
    """Astigmatism3DEstimation on the GPU equals the oracle (itself equal to the compiled reference)
    bit for bit, degenerate pairs included."""
    sig = _ast_inputs(20000, 5)
    got = pt_gpu.astig_estimate(sig, AST_PIXEL, model, zmax)
    exp = oracle_mod.astig_estimate(sig, AST_PIXEL, model, zmax)
    assert np.array_equal(got, exp, equal_nan=True), np.argwhere(~((got == exp) | (np.isnan(got) & np.isnan(exp))))[:3]


def test_astig_estimate_device_entry(pt_gpu, oracle_mod):
    """Device-pointer entry over 1M pairs: equal to the host entry, z inside the limits, scores in [0, 1]."""
    import torch
    n = 1 << 20
    rng = np.random.default_rng(9)
    z = rng.uniform(-600.0, 600.0, n)
    sig = _ast_pairs(AST_REF_MODEL, z) + rng.normal(0.0, 0.02, (n, 2))
    d_in = torch.from_numpy(sig).cuda()
    d_out = torch.empty_like(d_in)
    pt_gpu.astig_estimate_device(d_in, d_out, AST_PIXEL, AST_REF_MODEL, 600.0)
    got = d_out.cpu().numpy()
    assert np.all(np.abs(got[:, 0]) <= 600.0) and np.all((got[:, 1] >= 0) & (got[:, 1] <= 1))
    k = rng.choice(n, 4096, replace=False)
    assert np.array_equal(got[k], oracle_mod.astig_estimate(sig[k], AST_PIXEL, AST_REF_MODEL, 600.0))
    assert np.array_equal(got[:1000], pt_gpu.astig_estimate(sig[:1000], AST_PIXEL, AST_REF_MODEL, 600.0))


# ------------------------------------------------------------------ spline fit (mode 5) ----------
# The spline fit runs in FP64 with the reference's operations in its order (no contraction, correctly
# rounded sqrt and division), so its rows equal the oracle's (itself bit-exact to the compiled
# reference); the assertion allows the fit tolerance on the fitted columns for a device log ulp.
def assert_spline_rows(got, exp):
    assert got.shape == exp.shape and got.shape[0] > 0
    bad = fit_mismatches(got, exp)
    assert bad == [], bad[:5]
    same = np.all((got == exp) | (np.isnan(got) & np.isnan(exp)), axis=1)
    assert same.mean() >= 0.99, (same.mean(), fit_errors(got, exp))


@pytest.mark.parametrize("size,roi", [((11, 11, 21), 7), ((13, 13, 9), 9), ((9, 9, 5), 7), ((11, 11, 21), 8)])
def test_spline_localization_matches_oracle(pt_gpu, oracle_mod, size, roi):
    import synth_np
    from helpers import spline_params
    st = synth_np.frames(3, 96, 128, 14, seed=21)
    fp = spline_params(roi=roi, size=size, seed=5)
    thr = oracle_mod.autothreshold(st[0])
    got = pt_gpu.localization(st, thr, 0.0, 5, fp)
    exp = oracle_mod.localization(st, thr, 0.0, 5, fp)
    assert_spline_rows(got, exp)


def test_spline_fitting_matches_oracle(pt_gpu, oracle_mod, ref_vectors):
    from helpers import spline_params
    st = ref_vectors["dense"]
    pos = ref_vectors["dense_positions"]
    fp = spline_params()
    got = pt_gpu.fitting(st, pos, 5, fp)
    exp = oracle_mod.fitting(st, pos, 5, fp)
    assert np.array_equal(got[:, [ID, PLANE, INDEX]], exp[:, [ID, PLANE, INDEX]])
    assert_spline_rows(got, exp)


def test_spline_localization_512(pt_gpu, oracle_mod, frames512):
    """512 x 512 frames with the fusion: the spline rows of a chunked, multi-lane call."""
    from helpers import spline_params
    st, thr = frames512[0][:6], frames512[1]
    fp = spline_params(seed=3)
    got = pt_gpu.localization(st, thr, 10.0, 5, fp)
    exp = oracle_mod.localization(st, thr, 10.0, 5, fp)
    assert_spline_rows(got, exp)


def test_spline_bad_sizes_are_rejected(pt_gpu, capfd):
    """A spline size of 0 (where the reference reads out of bounds) returns no rows and says why."""
    st = np.full((1, 32, 32), 100, np.uint16)
    st[0, 16, 16] = 5000
    got = pt_gpu.localization(st, 10.0, 0.0, 5, [7.0, 0.0, 11.0, 11.0, 50.0])
    assert got.shape[0] == 0
    assert "spline sizes" in capfd.readouterr().err


# ------------------------------------------------------------------ large and dense frames ---------
@pytest.fixture(scope="module")
def large_frames(pt_gpu, oracle_mod):
    """1024^2 and 2048^2 frames at the headline density (200 emitters per 512^2) and a dense 512^2
    frame (1311 emitters, 0.5 per um^2 at 100 nm pixels): their exact frame floods hold more than
    16384 entries (the 2048^2 one more than 65536)."""
    import torch
    out = []
    for size, em, seed in [(1024, 800, 0x5EED2001), (512, 1311, 0x5EED2002), (2048, 3200, 0x5EED2003)]:
        d = torch.empty((1, size, size), dtype=torch.uint16, device="cuda")
        pt_gpu.synth_device(d, seed=seed, emitters=em)
        st = d.cpu().numpy()
        out.append((st, oracle_mod.autothreshold(st[0])))
    return out


@pytest.mark.parametrize("k", [0, 1, 2])
def test_large_frames_exact_flood(pt_gpu, oracle_mod, large_frames, k):
    """Every frame through the exact frame flood: labels equal the oracle's; the rank flood takes
    frames beyond its LDS sort (global-memory sort up to 65536 entries)."""
    st, thr = large_frames[k]
    w1 = oracle_mod.atrous(st[0])[1]
    exp = oracle_mod.watershed_labels(w1, thr)
    for forced in (0, 1):
        pt_gpu.set_option("force_sequential_watershed", forced)
        try:
            got = pt_gpu.debug_labels(st[0], thr, 0.0)
            stats = pt_gpu.last_stats()
        finally:
            pt_gpu.set_option("force_sequential_watershed", 0)
        assert stats["status"] == 0
        assert np.array_equal(rank_labels(got), exp), (forced, stats)
        if forced and k != 1:
            assert stats["rank_too_big"] == 0, stats


def test_large_frames_localization(pt_gpu, oracle_mod, large_frames):
    for st, thr in large_frames:
        got = pt_gpu.localization(st, thr, 0.0, 0, FIT_PARAMS_SYN)
        exp = oracle_mod.localization(st, thr, 0.0, 0, FIT_PARAMS_SYN)
        assert np.array_equal(got, exp)


def test_tuning_options_do_not_change_rows(pt_gpu, frames512):
    """Rows are identical with one or two main lanes and for any detection row-segment height."""
    st, thr = frames512
    base = pt_gpu.localization(st, thr, 10.0, 2, FIT_PARAMS_SYN)
    assert base.shape[0] > 0
    try:
        pt_gpu.set_option("chunk_pixels", 5 * 512 * 512)      # several chunks, so both lanes run
        for name, value in [("lanes", 1), ("lanes", 2), ("detect_rows", 37), ("detect_rows", 103),
                            ("detect_rows", 512)]:
            pt_gpu.set_option(name, value)
            got = pt_gpu.localization(st, thr, 10.0, 2, FIT_PARAMS_SYN)
            assert np.array_equal(got, base), (name, value)
    finally:
        pt_gpu.set_option("lanes", 2)
        pt_gpu.set_option("detect_rows", 0)
        pt_gpu.set_option("chunk_pixels", 0)
