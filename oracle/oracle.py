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
