"""Comparison helpers shared by the tests (no dependency on the library under test)."""
import numpy as np

# Column indices (lib/Constants.hpp:31-51).
ID, PLANE, INDEX, CHANNEL, X, Y, Z, INTENSITY, SX, SY, THETA, QXY, QZ, AMP, OFFSET, SEG, SURF, CIRC = range(18)


def theta_to_degrees(rows, fit):
    """The golden CSVs store theta in degrees in [-90, 90) for fit mode 4 (Tests_CPU/test_Palm.cpp:41-48)."""
    r = rows.copy()
    if fit == 4:
        m = (r[:, SX] > 0) & (r[:, SY] > 0)
        a = np.fmod(r[m, THETA] + np.pi / 2, np.pi)
        a[a < 0] += np.pi
        r[m, THETA] = (a - np.pi / 2) * 180 / np.pi
    return r


def pair_rows(actual, expected):
    """Index j(i) of the expected row paired with actual row i, by ExpectLocalizationsNear's rule
    (Tests_CPU/test_Palm.cpp:55-76): nearest unused expected row of the same plane and channel."""
    used = np.zeros(len(expected), bool)
    out = np.full(len(actual), -1, np.int64)
    for i, row in enumerate(actual):
        cand = np.where(~used & (expected[:, PLANE] == row[PLANE]) & (expected[:, CHANNEL] == row[CHANNEL]))[0]
        if cand.size == 0:
            continue
        d = ((expected[cand][:, [X, Y, Z]] - row[[X, Y, Z]]) ** 2).sum(1)
        j = cand[np.argmin(d)]
        used[j] = True
        out[i] = j
    return out


# Fit tolerance against the FP64 CPU path: positions within 1e-3 px; intensities, offset and sigmas
# within 1e-4 relative; theta within 1e-4 rad (relative beyond 1 rad); R^2 and circularity within
# 1e-4. Detection columns (Id, Plane, Index, Channel, Z, MSE Z, segmentation intensity, surface) are
# exact.
FIT_TOL = {X: (1e-3, 0.0), Y: (1e-3, 0.0), INTENSITY: (1e-9, 1e-4), AMP: (1e-9, 1e-4), OFFSET: (1e-9, 1e-4),
           SX: (1e-9, 1e-4), SY: (1e-9, 1e-4), THETA: (1e-4, 1e-4), QXY: (1e-4, 0.0), CIRC: (1e-4, 1e-4)}
EXACT_COLS = [ID, PLANE, INDEX, CHANNEL, Z, QZ, SEG, SURF]


def fit_mismatches(actual, expected):
    """Rows (paired as above) outside FIT_TOL, or whose fit failed on one side only."""
    assert actual.shape == expected.shape, (actual.shape, expected.shape)
    pair = pair_rows(actual, expected)
    bad = []
    for i, j in enumerate(pair):
        if j < 0:
            bad.append((i, "no candidate"))
            continue
        a, e = actual[i], expected[j]
        if (a[AMP] > 0) != (e[AMP] > 0):
            bad.append((i, "fit failed on one side", a[AMP], e[AMP]))
            continue
        for c in EXACT_COLS:
            if a[c] != e[c]:
                bad.append((i, c, a[c], e[c]))
        for c, (at, rt) in FIT_TOL.items():
            if not abs(a[c] - e[c]) <= at + rt * abs(e[c]):
                bad.append((i, c, a[c], e[c]))
    return bad


def fit_errors(actual, expected):
    """Largest absolute error per fit column over paired rows fitted on both sides (for reports)."""
    pair = pair_rows(actual, expected)
    ok = pair >= 0
    a, e = actual[ok], expected[pair[ok]]
    both = (a[:, AMP] > 0) & (e[:, AMP] > 0)
    a, e = a[both], e[both]
    if len(a) == 0:
        return {}
    names = {X: "x", Y: "y", INTENSITY: "intensity_rel", SX: "sigma_x_rel", SY: "sigma_y_rel", AMP: "amplitude_rel",
             OFFSET: "offset_rel"}
    out = {}
    for c, n in names.items():
        d = np.abs(a[:, c] - e[:, c])
        if n.endswith("_rel"):
            d = d / np.maximum(np.abs(e[:, c]), 1e-300)
        out[n] = float(d.max())
    return out


def match_rows(actual, expected, atol, rtol=1e-5, cols=None, col_tol=None):
    """The reference's ExpectLocalizationsNear (Tests_CPU/test_Palm.cpp:55-76): each actual row is
    paired with the nearest unused expected row of the same plane and channel (X, Y, Z distance),
    then every column must satisfy |a - e| <= atol + rtol * |e| (col_tol: per-column (atol, rtol)
    overrides). Returns the list of failures."""
    assert actual.shape == expected.shape, (actual.shape, expected.shape)
    used = np.zeros(len(expected), bool)
    bad = []
    cols = range(expected.shape[1]) if cols is None else cols
    for i, row in enumerate(actual):
        cand = np.where(~used & (expected[:, PLANE] == row[PLANE]) & (expected[:, CHANNEL] == row[CHANNEL]))[0]
        if cand.size == 0:
            bad.append((i, "no candidate"))
            continue
        d = ((expected[cand][:, [X, Y, Z]] - row[[X, Y, Z]]) ** 2).sum(1)
        j = cand[np.argmin(d)]
        used[j] = True
        e = expected[j]
        for c in cols:
            at, rt = (col_tol or {}).get(c, (atol, rtol))
            if not abs(row[c] - e[c]) <= at + rt * abs(e[c]):
                bad.append((i, c, row[c], e[c]))
    return bad


def rank_labels(labels):
    """Relabel a map whose labels are ordered like the reference's (any increasing identity) to 1..n."""
    u = np.unique(labels[labels > 0])
    lut = np.zeros(int(labels.max()) + 1 if labels.size else 1, np.uint32)
    lut[u] = np.arange(1, len(u) + 1, dtype=np.uint32)
    return lut[labels]


def filter_small(labels):
    """FilterData's surface filter (Segmentation2D.cpp:88-141) on an identity-labelled map."""
    u, c = np.unique(labels[labels > 0], return_counts=True)
    out = labels.copy()
    out[np.isin(out, u[c <= 4])] = 0
    return rank_labels(out)


def spline_params(roi=7, size=(11, 11, 21), delta_z=50.0, sigma=1.3, seed=0, noise=1e-3):
    """fit_params of the spline fit (fit mode 5): [roi, sx, sy, sz, deltaZ, coefficients], the
    coefficients of a separable Gaussian PSF interpolated per axis by cubic Hermite polynomials
    (lib/fit/SplineFit.cpp:36-49 layout: coefficient q of voxel v at q * sx*sy*sz + v, monomial
    x^(q % 4) y^((q / 4) % 4) z^(q / 16), v = (z * sy + y) * sx + x), with a small seeded
    perturbation so that no two coefficients are tied."""
    sx, sy, sz = size

    def axis(n, c, s):
        t = np.arange(n + 1, dtype=np.float64)
        g = np.exp(-(t - c) ** 2 / (2 * s * s))
        d = -(t - c) / (s * s) * g
        g0, g1, d0, d1 = g[:-1], g[1:], d[:-1], d[1:]
        return np.stack([g0, d0, 3 * (g1 - g0) - 2 * d0 - d1, 2 * (g0 - g1) + d0 + d1], axis=1)

    px = axis(sx, (sx - 1) / 2, sigma)
    py = axis(sy, (sy - 1) / 2, sigma)
    pz = axis(sz, (sz - 1) / 2, sz / 3)
    coef = np.einsum("za,yb,xc->abczyx", pz, py, px).reshape(64, sz * sy * sx)
    coef = coef * (1.0 + noise * np.random.default_rng(seed).standard_normal(coef.shape))
    return np.concatenate([[roi, sx, sy, sz, delta_z], coef.ravel()])


def pair_rows_by_plane(actual, expected):
    """Vectorized pairing for large row sets: within each plane, each actual row takes the expected
    row nearest in (X, Y). Returns the index array (-1 where the plane has no expected rows)."""
    out = np.full(len(actual), -1, np.int64)
    order_a = np.argsort(actual[:, PLANE], kind="stable")
    order_e = np.argsort(expected[:, PLANE], kind="stable")
    pa, pe = actual[order_a, PLANE], expected[order_e, PLANE]
    for plane in np.unique(pa):
        ia = order_a[np.searchsorted(pa, plane, "left"):np.searchsorted(pa, plane, "right")]
        ie = order_e[np.searchsorted(pe, plane, "left"):np.searchsorted(pe, plane, "right")]
        if ie.size == 0:
            continue
        d = ((actual[ia][:, None, X:Y + 1] - expected[ie][None, :, X:Y + 1]) ** 2).sum(-1)
        out[ia] = ie[d.argmin(1)]
    return out


def fit_mismatch_rows(actual, expected):
    """Indices of actual rows outside FIT_TOL of their expected row (paired by pair_rows_by_plane), with
    unequal detection columns, fitted on one side only, unpaired, or paired twice. For row sets of
    tens of thousands of rows."""
    assert actual.shape == expected.shape, (actual.shape, expected.shape)
    j = pair_rows_by_plane(actual, expected)
    bad = j < 0
    jj = np.where(bad, 0, j)
    e = expected[jj]
    dup = np.zeros(len(actual), bool)
    u, c = np.unique(j[j >= 0], return_counts=True)
    dup[np.isin(j, u[c > 1])] = True
    bad |= dup
    bad |= (actual[:, AMP] > 0) != (e[:, AMP] > 0)
    for c in EXACT_COLS:
        if c in (ID, INDEX):
            continue                   # Id / Index follow the (Y, X) order, checked through the pairing
        bad |= actual[:, c] != e[:, c]
    for c, (at, rt) in FIT_TOL.items():
        bad |= ~(np.abs(actual[:, c] - e[:, c]) <= at + rt * np.abs(e[:, c]))
    return np.nonzero(bad)[0]
