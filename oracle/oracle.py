"""ctypes access to the CPU oracle (liboracle.so) and, where built, the compiled reference (_ref/libref.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
leg, to compare results against; the product never imports it.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(HERE, "liboracle.so")
REF_SO = os.path.join(HERE, "_ref", "libref.so")
NCOL = 18
_u64, _f64, _sz, _vp, _i32 = ctypes.c_uint64, ctypes.c_double, ctypes.c_size_t, ctypes.c_void_p, ctypes.c_int
_or = None
_ref = None


def build():
    """Compile the oracle (and the reference when its tree is mounted)."""
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _or
    if _or is None:
        if not os.path.exists(ORACLE_SO):
            build()
        L = ctypes.CDLL(ORACLE_SO)
        L.or_atrous.argtypes = [_vp, _sz, _sz, _vp, _vp, _vp]
        L.or_segment.restype = _sz
        L.or_segment.argtypes = [_vp, _sz, _sz, _f64, _f64, _vp, _vp, _sz]
        L.or_watershed_labels.restype = _sz
        L.or_watershed_labels.argtypes = [_vp, _sz, _sz, _f64, _vp]
        L.or_fit_region.argtypes = [_i32, _vp, _sz, _sz, _vp, _i32, _vp, _vp]
        L.or_inverse_gauss_jordan.restype = _i32
        L.or_inverse_gauss_jordan.argtypes = [_vp, _vp, _sz]
        L.or_localization.restype = _u64
        L.or_localization.argtypes = [_vp, _vp, _u64, _u64, _u64, _u64, _f64, _f64, _u64, _vp, _i32]
        L.or_fitting.restype = _u64
        L.or_fitting.argtypes = [_vp, _vp, _u64, _vp, _u64, _u64, _u64, _u64, _u64, _vp, _i32]
        L.or_wavelett.argtypes = [_vp, _vp, _u64, _u64, _u64, _u64]
        L.or_autothreshold.restype = _f64
        L.or_autothreshold.argtypes = [_vp, _u64, _u64, _vp]
        L.or_astig_estimate.argtypes = [_vp, _vp, _u64, _f64, _vp, _f64]
        L.or_astig_quality.restype = _f64
        L.or_astig_quality.argtypes = [_vp, _f64, _f64, _f64, _f64]
        L.or_libm.restype = None
        L.or_libm.argtypes = [_i32, _vp, _u64, _vp]
        _or = L
    return _or


def ref_lib():
    """The compiled reference, or None when it was not built (no reference tree here)."""
    global _ref
    if _ref is None and os.path.exists(REF_SO):
        L = ctypes.CDLL(REF_SO)
        L.ref_localization.restype = _u64
        L.ref_localization.argtypes = [_vp, _vp, _u64, _u64, _u64, _u64, _f64, _f64, _u64, _vp]
        L.ref_fitting.restype = _u64
        L.ref_fitting.argtypes = [_vp, _vp, _u64, _vp, _u64, _u64, _u64, _u64, _u64, _vp]
        L.ref_wavelett.argtypes = [_vp, _vp, _u64, _u64, _u64, _u64]
        L.ref_autothreshold.restype = _f64
        L.ref_autothreshold.argtypes = [_vp, _u64, _u64, _vp]
        L.ref_segment.restype = _u64
        L.ref_segment.argtypes = [_vp, _u64, _u64, _f64, _f64, _vp, _vp, _u64]
        L.ref_atrous.argtypes = [_vp, _u64, _u64, _vp, _vp, _vp]
        L.ref_astig_estimate.argtypes = [_vp, _vp, _u64, _f64, _vp, _f64]
        _ref = L
    return _ref


def _stack(s):
    s = np.ascontiguousarray(s, np.uint16)
    return s[None] if s.ndim == 2 else s


def _params(fp):
    """fit_params as the exports read them: at least 4 values (zero-padded); longer arrays (the spline
    fit's [roi, sizes, deltaZ, coefficients]) are passed whole."""
    a = np.asarray(fp, np.float64).ravel()
    p = np.zeros(max(4, a.size))
    p[:a.size] = a
    return p


def default_capacity(z, h, w):
    return ((h * w) // 5 + 1) * z * NCOL


def atrous(frame):
    f = np.ascontiguousarray(frame, np.uint16)
    h, w = f.shape
    planes = [np.empty((h, w)) for _ in range(3)]
    lib().or_atrous(f.ctypes.data, w, h, *[p.ctypes.data for p in planes])
    return planes


def watershed_labels(w1, threshold):
    w1 = np.ascontiguousarray(w1, np.float64)
    h, w = w1.shape
    out = np.zeros((h, w), np.uint32)
    lib().or_watershed_labels(w1.ctypes.data, w, h, float(threshold), out.ctypes.data)
    return out


def segment(w1, threshold, ratio=0.0):
    """Final label map and regions [x, y, z, intensity, surface] in label order."""
    w1 = np.ascontiguousarray(w1, np.float64)
    h, w = w1.shape
    labels = np.zeros((h, w), np.uint32)
    cap = h * w // 4 + 2
    reg = np.zeros((cap, 5))
    n = lib().or_segment(w1.ctypes.data, w, h, float(threshold), float(ratio), labels.ctypes.data, reg.ctypes.data, cap)
    return labels, reg[:n].copy()


def localization(stack, threshold, ratio=0.0, fit=2, fit_params=(7, 1.3, 1.3, 0), output_size=None, max_iter=75):
    s = _stack(stack)
    z, h, w = s.shape
    cap = default_capacity(z, h, w) if output_size is None else int(output_size)
    out = np.empty(max(cap, 1))
    p = _params(fit_params)
    n = lib().or_localization(s.ctypes.data, out.ctypes.data, cap, h, w, z, float(threshold), float(ratio), int(fit),
                              p.ctypes.data, int(max_iter))
    return out[:n].reshape(-1, NCOL).copy()


def fitting(stack, positions, fit=2, fit_params=(7, 1.3, 1.3, 0), max_iter=75):
    s = _stack(stack)
    z, h, w = s.shape
    pos = np.ascontiguousarray(positions, np.float64).reshape(-1, 6)
    out = np.empty(max(pos.shape[0] * NCOL, 1))
    p = _params(fit_params)
    n = lib().or_fitting(s.ctypes.data, pos.ctypes.data, pos.shape[0], out.ctypes.data, out.size, h, w, z, int(fit),
                         p.ctypes.data, int(max_iter))
    return out[:n].reshape(-1, NCOL).copy()


def wavelett(stack, level=2):
    s = _stack(stack)
    z, h, w = s.shape
    out = np.empty((z, h, w))
    lib().or_wavelett(s.ctypes.data, out.ctypes.data, h, w, z, int(level))
    return out


def autothreshold(frame, fit_params=(7,)):
    f = np.ascontiguousarray(frame, np.uint16)
    h, w = f.shape
    p = _params(fit_params)
    return lib().or_autothreshold(f.ctypes.data, h, w, p.ctypes.data)


def inverse_gauss_jordan(a):
    a = np.array(a, np.float64, copy=True, order="C")
    n = a.shape[0]
    out = np.zeros_like(a)
    ok = lib().or_inverse_gauss_jordan(a.ctypes.data, out.ctypes.data, n) == 0
    return ok, out


def fit_region(mode, frame, fit_params, x, y, seg=0.0, surf=0.0, max_iter=75):
    f = np.ascontiguousarray(frame, np.uint16)
    h, w = f.shape
    info = np.array([x, y, 0.0, seg, surf])
    row = np.zeros(NCOL)
    p = _params(fit_params)
    lib().or_fit_region(int(mode), f.ctypes.data, w, h, p.ctypes.data, int(max_iter), info.ctypes.data, row.ctypes.data)
    return row


def astig_estimate(sigmas, pixel_size, model, z_max):
    """AstigFit::Estimate restated: (n, 2) sigma pairs in pixels -> (n, 2) [z nm, axial score]."""
    s = np.ascontiguousarray(sigmas, np.float64).reshape(-1, 2)
    m = np.ascontiguousarray(model, np.float64).ravel()
    assert m.size == 10
    out = np.empty_like(s)
    lib().or_astig_estimate(s.ctypes.data, out.ctypes.data, s.shape[0], float(pixel_size), m.ctypes.data, float(z_max))
    return out


def astig_quality(model, z, sx, sy, pixel_size):
    """AstigFit::getQuality restated."""
    m = np.ascontiguousarray(model, np.float64).ravel()
    return lib().or_astig_quality(m.ctypes.data, float(z), float(sx), float(sy), float(pixel_size))


def libm(name, x):
    """exp, sin or cos of the C library the oracle's fits call, elementwise."""
    fn = {"exp": 0, "sin": 1, "cos": 2}[name]
    x = np.ascontiguousarray(x, np.float64)
    y = np.empty_like(x)
    lib().or_libm(fn, x.ctypes.data, x.size, y.ctypes.data)
    return y


def ref_astig_estimate(sigmas, pixel_size, model, z_max):
    """The compiled reference's AstigFit::Estimate (None without oracle/_ref)."""
    R = ref_lib()
    if R is None:
        return None
    s = np.ascontiguousarray(sigmas, np.float64).reshape(-1, 2)
    m = np.ascontiguousarray(model, np.float64).ravel()
    out = np.empty_like(s)
    R.ref_astig_estimate(s.ctypes.data, out.ctypes.data, s.shape[0], float(pixel_size), m.ctypes.data, float(z_max))
    return out


def ref_localization(stack, threshold, ratio=0.0, fit=2, fit_params=(7, 1.3, 1.3, 0), output_size=None):
    """The compiled reference's Localization (None without oracle/_ref)."""
    R = ref_lib()
    if R is None:
        return None
    s = _stack(stack)
    z, h, w = s.shape
    cap = default_capacity(z, h, w) if output_size is None else int(output_size)
    out = np.empty(max(cap, 1))
    p = _params(fit_params)
    n = R.ref_localization(s.ctypes.data, out.ctypes.data, cap, h, w, z, float(threshold), float(ratio), int(fit),
                           p.ctypes.data)
    return out[:n].reshape(-1, NCOL).copy()


def ref_fitting(stack, positions, fit=2, fit_params=(7, 1.3, 1.3, 0)):
    """The compiled reference's Fitting (None without oracle/_ref)."""
    R = ref_lib()
    if R is None:
        return None
    s = _stack(stack)
    z, h, w = s.shape
    pos = np.ascontiguousarray(positions, np.float64).reshape(-1, 6)
    out = np.empty(max(pos.shape[0] * NCOL, 1))
    p = _params(fit_params)
    n = R.ref_fitting(s.ctypes.data, pos.ctypes.data, pos.shape[0], out.ctypes.data, out.size, h, w, z, int(fit),
                      p.ctypes.data)
    return out[:n].reshape(-1, NCOL).copy()


def gauss_model(mode, p, dx, dy, sigma, theta):
    """Gaussian model of the fits (lib/fit/GaussianModel.cpp:19-52) over pixel offsets dx, dy from the
    centre, parameters p = [offset, amplitude, x, y, (sigma | sqrt2 sigmaX, sqrt2 sigmaY, (theta))]
    (the centre enters through dx, dy: callers pass pixel - p[2], pixel - p[3])."""
    if mode <= 2:
        s = sigma if mode == 1 else p[4]
        return p[0] + p[1] * np.exp(-(dx * dx + dy * dy) / (2.0 * s * s))
    th = theta if mode == 3 else p[6]
    c, s = np.cos(th), np.sin(th)
    ix2, iy2 = 1.0 / (p[4] * p[4]), 1.0 / (p[5] * p[5])
    ax, ay, axy = c * c * ix2 + s * s * iy2, s * s * ix2 + c * c * iy2, 2.0 * c * s * (iy2 - ix2)
    return p[0] + p[1] * np.exp(-(ax * dx * dx + ay * dy * dy + axy * dx * dy))


def crlb(rows, height, width, fit, fit_params, camera=(1.0, 0.0, 0.0)):
    """Cramer-Rao bounds of Gaussian-fit rows (the quantities of the library's pt_crlb_device): the
    Fisher information sum g g^T / var over the fit-side ROI centred on the fitted pixel, with g the
    model gradient taken here by central finite differences (not the library's analytic gradient),
    var = gain * (model - offset) + readVar; bounds sqrt(diag(F^-1)) mapped to the row's columns
    [X, Y, SigmaX, SigmaY, Theta, Amplitude, Offset, Intensity], intensity by the delta method."""
    rows = np.asarray(rows, np.float64).reshape(-1, NCOL)
    roi, sig, _, th = float(fit_params[0]), float(fit_params[1]), float(fit_params[2]), float(fit_params[3])
    half = int(roi) // 2
    gain, off_cam, rvar = camera
    npar = {1: 4, 2: 5, 3: 6, 4: 7}[fit]
    out = np.full((len(rows), 8), np.nan)
    ii, jj = np.meshgrid(np.arange(-half, half + 1), np.arange(-half, half + 1), indexing="ij")
    for r, row in enumerate(rows):
        X, Y, amp, off = row[4], row[5], row[13], row[14]
        if not (amp > 0) or not np.isfinite(X) or not np.isfinite(Y):
            continue
        fx, fy = np.floor(X), np.floor(Y)
        if fx - half < 0 or fx + half >= width or fy - half < 0 or fy + half >= height:
            continue
        p = np.array([off, amp, X, Y, 0.0, 0.0, 0.0][:npar])
        if fit == 2:
            p[4] = row[8]
        if fit >= 3:
            p[4], p[5] = np.sqrt(2.0) * row[8], np.sqrt(2.0) * row[9]
        if fit == 4:
            p[6] = row[10]
        px, py = fx + jj, fy + ii

        def model(q):
            return gauss_model(fit, q, px - q[2], py - q[3], sig, th)

        g = np.empty((npar,) + px.shape)
        for k in range(npar):
            h = 1e-6 * max(1.0, abs(p[k]))
            a, b = p.copy(), p.copy()
            a[k] += h
            b[k] -= h
            g[k] = (model(a) - model(b)) / (2 * h)
        circ = fit == 4 and abs(p[4] - p[5]) <= 2.220446049250313e-16
        if circ:
            g[6] = 0.0
        var = np.maximum(gain * np.maximum(model(p) - off_cam, 0.0) + rvar, 1e-12)
        F = np.einsum("aij,bij->ab", g, g / var)
        if circ:
            F[6, 6] = 1.0
        try:
            np.linalg.cholesky(F)
        except np.linalg.LinAlgError:
            continue
        Fi = np.linalg.inv(F)
        o = out[r]
        o[0], o[1], o[5], o[6] = np.sqrt(Fi[2, 2]), np.sqrt(Fi[3, 3]), np.sqrt(Fi[1, 1]), np.sqrt(Fi[0, 0])
        d = np.zeros(npar)
        if fit == 1:
            d[1] = 2 * np.pi * sig * sig
        if fit == 2:
            o[2] = o[3] = np.sqrt(Fi[4, 4])
            d[1], d[4] = 2 * np.pi * p[4] ** 2, 4 * np.pi * amp * p[4]
        if fit >= 3:
            o[2], o[3] = np.sqrt(Fi[4, 4]) / np.sqrt(2.0), np.sqrt(Fi[5, 5]) / np.sqrt(2.0)
            d[1], d[4], d[5] = np.pi * p[4] * p[5], np.pi * amp * p[5], np.pi * amp * p[4]
        if fit == 4:
            o[4] = np.nan if circ else np.sqrt(Fi[6, 6])
        o[7] = np.sqrt(max(d @ Fi @ d, 0.0))
    return out


def mle_fit(frame, fit, fit_params, x, y, camera=(1.0, 0.0, 0.0)):
    """Poisson maximum-likelihood fit of one ROI (the library's fit objective 1): the reference's ROI,
    border rule, initial values and bounds (GaussianFit.cpp:51-56, Fitting.cpp:126-175), the
    objective sum mu - d log mu with mu = (model - offset) / gain + rv, d = max(0, (v - offset) /
    gain) + rv, rv = readVar / gain^2, minimised by scipy's bounded L-BFGS-B with numerical gradients
    (an optimiser independent of the library's Fisher-scoring LM). Returns [offset, amplitude, x, y,
    sigmaX, sigmaY, theta] in the row's units, or None for the border case."""
    from scipy.optimize import minimize
    img = np.asarray(frame, np.float64)
    H, W = img.shape
    roi, sig, sigy, th = int(fit_params[0]), float(fit_params[1]), float(fit_params[2]), float(fit_params[3])
    if roi % 2 == 0:
        roi -= 1
    half = roi // 2
    ux, uy = int(x), int(y)
    if half > W or half > H or half > ux or ux >= W - half or half > uy or uy >= H - half:
        return None
    patch = img[uy - half:uy + half + 1, ux - half:ux + half + 1]
    gain, c0, rvar = camera
    rv = rvar / (gain * gain)
    yy, xx = np.mgrid[uy - half:uy + half + 1, ux - half:ux + half + 1].astype(np.float64)
    d = np.maximum(0.0, (patch - c0) / gain) + rv
    p = [patch.min(), patch.max() - patch.min(), x, y]
    lo, hi = [0.0, 0.0, x - 1, y - 1], [65535.0, 65535.0, x + 1, y + 1]
    if fit == 2:
        p += [sig]; lo += [0.1]; hi += [20.0]
    if fit >= 3:
        p += [np.sqrt(2) * sig, np.sqrt(2) * sigy]; lo += [0.1, 0.1]; hi += [20.0, 20.0]
    if fit == 4:
        p += [th]; lo += [-150.0]; hi += [150.0]
    p = np.array(p)
    if p[0] <= c0:
        p[0] = c0 + 1e-3 * gain
    # The likelihood is defined where every mu > 0; with amplitude >= 0 that is offset > c0 - rv gain,
    # so the box keeps the optimiser inside the domain (the library rejects steps that leave it).
    lo[0] = max(lo[0], c0 - rv * gain + 1e-6 * gain)

    # The Poisson deviance sum mu - d - d log(mu / d): the same minimiser as sum mu - d log mu, but of
    # the order of the pixel count, so that finite-difference gradients keep their precision.
    dlogd = np.where(d > 0, d * np.log(np.where(d > 0, d, 1.0)), 0.0)

    def nll(q):
        mu = (gauss_model(fit, q, xx - q[2], yy - q[3], sig, th) - c0) / gain + rv
        if np.any(mu <= 0):
            return 1e300
        return float(np.sum(mu - d - np.where(d > 0, d * np.log(mu), 0.0) + dlogd))

    scale = np.maximum(np.abs(p), 1.0)
    bounds = list(zip(np.array(lo) / scale, np.array(hi) / scale))
    z = p / scale
    for _ in range(3):  # restarts from the last point until L-BFGS-B stops moving
        res = minimize(lambda v: nll(v * scale), z, method="L-BFGS-B", jac="3-point", bounds=bounds,
                       options={"ftol": 0.0, "gtol": 1e-9, "maxiter": 3000})
        moved = np.abs(res.x - z).max()
        z = res.x
        if moved < 1e-9:
            break
    q = z * scale
    sx = sy = sig
    t = th
    if fit == 1:
        t = 0.0
    if fit == 2:
        sx = sy = q[4]
        t = 0.0
    if fit >= 3:
        sx, sy = q[4] / np.sqrt(2), q[5] / np.sqrt(2)
    if fit == 4:
        t = q[6]
    return np.array([q[0], q[1], q[2], q[3], sx, sy, t])
