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


def match_rows(actual, expected, atol, rtol=1e-5, cols=None):
    """The reference's ExpectLocalizationsNear (Tests_CPU/test_Palm.cpp:55-76): each actual row is
    paired with the nearest unused expected row of the same plane and channel (X, Y, Z distance),
    then every column must satisfy |a - e| <= atol + rtol * |e|. Returns the list of failures."""
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
            if not abs(row[c] - e[c]) <= atol + rtol * abs(e[c]):
                bad.append((i, c, row[c], e[c]))
    return bad


def rank_labels(labels):
    """Relabel a map whose labels are ordered like the reference's (any increasing identity) to 1..n."""
    u = np.unique(labels[labels > 0])
    lut = np.zeros(int(labels.max()) + 1 if labels.size else 1, np.uint32)
    lut[u] = np.arange(1, len(u) + 1, dtype=np.uint32)
    return lut[labels]
