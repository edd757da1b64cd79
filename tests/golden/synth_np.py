"""Seeded synthetic single-molecule frames in numpy (inputs of the committed fixtures).

ADU = 100 + 2 * Poisson(20 + sum_k N_k * PSF_k), PSF_k the pixel-integrated Gaussian (sigma 1.3 px),
N_k ~ LogNormal(ln 2000, 0.5), positions uniform: the small-stack configuration of the benchmark.
The generated frames are stored in the fixtures, so their exact values never depend on numpy.
"""
import numpy as np
from scipy.special import erf


def frames(n, height, width, emitters, seed, sigma=1.3, bg=20.0, median=2000.0, sigma_log=0.5, offset=100, gain=2):
    rng = np.random.default_rng(seed)
    out = np.empty((n, height, width), np.uint16)
    r = 7
    s2 = np.sqrt(2.0) * sigma
    for f in range(n):
        lam = np.full((height, width), bg)
        xs = rng.uniform(0, width, emitters)
        ys = rng.uniform(0, height, emitters)
        ns = np.exp(np.log(median) + sigma_log * rng.standard_normal(emitters))
        for x, y, k in zip(xs, ys, ns):
            x0, x1 = max(0, int(x) - r), min(width, int(x) + r + 1)
            y0, y1 = max(0, int(y) - r), min(height, int(y) + r + 1)
            px, py = np.arange(x0, x1), np.arange(y0, y1)
            gx = 0.5 * (erf((px + 0.5 - x) / s2) - erf((px - 0.5 - x) / s2))
            gy = 0.5 * (erf((py + 0.5 - y) / s2) - erf((py - 0.5 - y) / s2))
            lam[y0:y1, x0:x1] += k * np.outer(gy, gx)
        out[f] = np.minimum(offset + gain * rng.poisson(lam), 65535)
    return out
