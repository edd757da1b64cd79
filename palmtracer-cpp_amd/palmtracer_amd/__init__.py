"""ctypes binding of libpalmtracer_amd.so, shaped like the reference's Python front end.

The reference library (tmonseigne/PALMTracer-cpp) is driven from Python through ctypes with host
numpy buffers (CPU/main.cpp comments :106-107, :229-230). The functions below call the same C entry
points with the same argument meaning; `*_device` variants take torch tensors that live in HBM.

The library must have been built (palmtracer-cpp_amd/build.py); there is no CPU fallback.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# The in-tree library; PALMTRACER_AMD_LIBRARY names another build of it.
LIB_PATH = os.environ.get("PALMTRACER_AMD_LIBRARY") or os.path.join(os.path.dirname(HERE), "libpalmtracer_amd.so")
COLUMNS = ["Id", "Plane", "Index", "Channel", "X", "Y", "Z", "Integrated Intensity", "Sigma X", "Sigma Y", "Theta",
           "MSE XY", "MSE Z", "Intensity 0", "Intensity Offset", "Intensity", "Surface", "Circularity"]
NCOL = 18

_u64, _f64, _i32, _vp = ctypes.c_uint64, ctypes.c_double, ctypes.c_int, ctypes.c_void_p
_lib = None

# Default synthetic camera (single-molecule stacks of the benchmark configurations).
SYNTH_2D = dict(bg=20.0, sigma=1.3, median=2000.0, sigma_log=0.5, offset=100.0, gain=2.0, astig=0, s0=1.2, c=250.0,
                d=400.0, zmin=-600.0, zmax=600.0, scmos=0, sd_offset=3.0, sd_gain=0.1)


class BatchStack(ctypes.Structure):
    """pt_batch_stack of include/palmtracer.h."""
    _fields_ = [("d_input", _vp), ("planes", _u64), ("height", _u64), ("width", _u64), ("thresholdVal", _f64),
                ("watershedRatio", _f64), ("fit", _u64), ("fit_params", _vp), ("d_output", _vp),
                ("outputCapDoubles", _u64), ("maxPerPlane", _u64), ("doublesWritten", _u64)]


def lib():
    """Load the shared library once; raise if it was not built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: run palmtracer-cpp_amd/build.py (no CPU fallback exists)")
        L = ctypes.CDLL(LIB_PATH)
        L.Localization.restype = _u64
        L.Localization.argtypes = [_vp, _vp, _u64, _u64, _u64, _u64, _f64, _f64, _u64, _vp]
        L.Fitting.restype = _u64
        L.Fitting.argtypes = [_vp, _vp, _u64, _vp, _u64, _u64, _u64, _u64, _u64, _vp]
        L.Wavelett.restype = None
        L.Wavelett.argtypes = [_vp, _vp, _u64, _u64, _u64, _u64]
        L.pt_localization_device.restype = _u64
        L.pt_localization_device.argtypes = [_vp, _vp, _u64, _u64, _u64, _u64, _f64, _f64, _u64, _vp, _u64]
        L.pt_wavelet_device.restype = _i32
        L.pt_wavelet_device.argtypes = [_vp, _vp, _u64, _u64, _u64, _u64]
        L.pt_debug_labels.restype = _i32
        L.pt_debug_labels.argtypes = [_vp, _u64, _u64, _f64, _f64, _vp]
        L.pt_debug_libm.restype = _i32
        L.pt_debug_libm.argtypes = [_u64, _vp, _vp, _u64]
        L.AutoThreshold.restype = _f64
        L.AutoThreshold.argtypes = [_vp, _u64, _u64, _vp]
        L.pt_autothreshold_device.restype = _f64
        L.pt_autothreshold_device.argtypes = [_vp, _u64, _u64, _vp]
        L.pt_set_option.restype = _i32
        L.pt_set_option.argtypes = [ctypes.c_char_p, _f64]
        L.pt_stage_times.restype = _i32
        L.pt_stage_times.argtypes = [_vp, _u64]
        L.pt_last_stats.restype = _i32
        L.pt_last_stats.argtypes = [_vp, _u64]
        L.pt_synth_device.restype = _i32
        L.pt_synth_device.argtypes = [_vp, _u64, _i32, _i32, _i32, _i32, _vp]
        L.pt_synth_emitters_device.restype = _i32
        L.pt_synth_emitters_device.argtypes = [_vp, _u64, _i32, _i32, _i32, _i32, _vp]
        L.Astigmatism3DEstimation.restype = None
        L.Astigmatism3DEstimation.argtypes = [_vp, _vp, _u64, _f64, _vp, _f64]
        L.pt_astig_estimate_device.restype = _i32
        L.pt_astig_estimate_device.argtypes = [_vp, _vp, _u64, _f64, _vp, _f64]
        L.pt_fitting_noalloc.restype = _u64
        L.pt_fitting_noalloc.argtypes = [_vp, _vp, _u64, _vp, _u64, _u64, _u64, _u64, _vp]
        L.pt_crlb_device.restype = _i32
        L.pt_crlb_device.argtypes = [_vp, _u64, _u64, _u64, _u64, _vp, _vp, _vp]
        L.pt_batch_localization.restype = _i32
        L.pt_batch_localization.argtypes = [ctypes.POINTER(BatchStack), _u64]
        L.pt_batch_schedules.restype = _i32
        L.pt_batch_schedules.argtypes = [_vp, _u64]
        L.pt_server_create.restype = _vp
        L.pt_server_create.argtypes = [_u64, _u64, _u64]
        L.pt_server_submit.restype = ctypes.c_int64
        L.pt_server_submit.argtypes = [_vp, _vp, _u64, _u64, _u64, _f64, _f64, _u64, _vp, _u64, _vp, _u64, _u64]
        L.pt_server_wait.restype = _i32
        L.pt_server_wait.argtypes = [_vp, ctypes.c_int64, ctypes.POINTER(ctypes.c_uint64), ctypes.POINTER(ctypes.c_double)]
        L.pt_server_request_bytes.restype = _u64
        L.pt_server_request_bytes.argtypes = [_vp, _u64, _u64, _u64, _u64, _u64]
        L.pt_server_now.restype = _f64
        L.pt_server_now.argtypes = [_vp]
        L.pt_server_destroy.restype = None
        L.pt_server_destroy.argtypes = [_vp]
        _lib = L
    return _lib


def _stack(stack):
    s = np.ascontiguousarray(stack, dtype=np.uint16)
    if s.ndim == 2:
        s = s[None]
    if s.ndim != 3:
        raise ValueError("stack must be (planes, height, width) or (height, width)")
    return s


def _params(fit_params):
    """fit_params as a contiguous float64 array of at least 4 values (the Gaussian modes read 4; the
    spline mode 5 reads [roi, sx, sy, sz, deltaZ] and 64 * sx * sy * sz coefficients)."""
    fp = np.asarray(fit_params, np.float64).ravel()
    p = np.zeros(max(4, fp.size), np.float64)
    p[:fp.size] = fp
    return p


def default_capacity(planes, height, width):
    """Output size (doubles) the reference's tests allocate: ((h*w)/5 + 1) * planes * 18."""
    return ((height * width) // 5 + 1) * planes * NCOL


def localization(stack, threshold, watershed_ratio=0.0, fit=2, fit_params=(7, 1.3, 1.3, 0), output_size=None):
    """Localization (CPU/main.cpp:112-121): rows (n, 18) in the reference's column order."""
    s = _stack(stack)
    z, h, w = s.shape
    cap = default_capacity(z, h, w) if output_size is None else int(output_size)
    out = np.empty(max(cap, 1), np.float64)
    p = _params(fit_params)
    n = lib().Localization(s.ctypes.data, out.ctypes.data, cap, h, w, z, float(threshold), float(watershed_ratio),
                           int(fit), p.ctypes.data)
    return out[:n].reshape(-1, NCOL).copy()


def fitting(stack, positions, fit=2, fit_params=(7, 1.3, 1.3, 0)):
    """Fitting (CPU/main.cpp:141-148): positions (n, 6) = [Id, Plane, X, Y, SegIntensity, Surface]."""
    s = _stack(stack)
    z, h, w = s.shape
    pos = np.ascontiguousarray(positions, np.float64).reshape(-1, 6)
    out = np.empty(max(pos.shape[0] * NCOL, 1), np.float64)
    p = _params(fit_params)
    n = lib().Fitting(s.ctypes.data, pos.ctypes.data, pos.shape[0], out.ctypes.data, out.size, h, w, z, int(fit),
                      p.ctypes.data)
    return out[:n].reshape(-1, NCOL).copy()


def wavelett(stack, level=2):
    """Wavelett (CPU/main.cpp:257-284): one plane per frame, level 1 = W0, 2 = W1, 3 = V2."""
    s = _stack(stack)
    z, h, w = s.shape
    out = np.empty((z, h, w), np.float64)
    lib().Wavelett(s.ctypes.data, out.ctypes.data, h, w, z, int(level))
    return out


def autothreshold(frame, fit_params=(7,)):
    """AutoThreshold (CPU/main.cpp:75-85) of one plane."""
    f = np.ascontiguousarray(frame, np.uint16)
    h, w = f.shape
    p = _params(fit_params)
    return lib().AutoThreshold(f.ctypes.data, h, w, p.ctypes.data)


def autothreshold_device(d_frame, fit_params=(7,)):
    h, w = d_frame.shape[-2:]
    p = _params(fit_params)
    return lib().pt_autothreshold_device(d_frame.data_ptr(), h, w, p.ctypes.data)


def debug_labels(frame, threshold, watershed_ratio=0.0):
    f = np.ascontiguousarray(frame, np.uint16)
    h, w = f.shape
    out = np.zeros((h, w), np.uint32)
    if not lib().pt_debug_labels(f.ctypes.data, h, w, float(threshold), float(watershed_ratio), out.ctypes.data):
        raise RuntimeError("pt_debug_labels failed")
    return out


def debug_libm(name, d_x, d_y):
    """d_y = f(d_x) on the device with the fit model's exp, sin or cos (float64 tensors of one size)."""
    fn = {"exp": 0, "sin": 1, "cos": 2}[name]
    assert d_x.dtype == d_y.dtype and d_x.numel() == d_y.numel() and d_x.is_contiguous() and d_y.is_contiguous()
    if not lib().pt_debug_libm(fn, d_x.data_ptr(), d_y.data_ptr(), d_x.numel()):
        raise RuntimeError("pt_debug_libm failed")


def set_option(name, value):
    if not lib().pt_set_option(name.encode(), float(value)):
        raise ValueError(f"option {name!r} rejected")


def last_stats():
    v = (ctypes.c_uint64 * 24)()
    lib().pt_last_stats(v, 24)
    keys = ["frames", "chunks", "rows", "flagged_frames", "multi_components", "deferred_frames", "status",
            "flag_overflow", "flag_unproven", "wave_flood_pops", "wave_flood_max_queue", "rank_too_big",
            "rank_equal_keys", "rank_unsettled", "wave_queue_overflow", "heap_frames", "rank_two_orders",
            "fit_handed_over", "grid_frames", "grid_equal_keys", "grid_unsettled", "grid_rejected", "grid_other",
            "mask_pixels"]
    return dict(zip(keys, list(v)))


def stage_times():
    """Per-stage device milliseconds of the last Localization (needs set_option('profile', 1))."""
    v = (ctypes.c_double * 13)()
    lib().pt_stage_times(v, 13)
    keys = ["atrous_ms", "segment_ms", "fit_ms", "assemble_ms", "chunks", "jacobian_passes", "trial_passes",
            "deferred_ms", "group_jacobian_passes", "group_trial_passes", "h2d_ms", "d2h_ms", "stream_pieces"]
    return dict(zip(keys, list(v)))


def _synth_params(cfg):
    c = dict(SYNTH_2D)
    c.update(cfg or {})
    order = ["bg", "sigma", "median", "sigma_log", "offset", "gain", "astig", "s0", "c", "d", "zmin", "zmax", "scmos",
             "sd_offset", "sd_gain"]
    return np.array([float(c[k]) for k in order], np.float64)


def synth_device(out, seed, emitters, **cfg):
    """Fill the torch uint16 tensor `out` (frames, height, width) on the GPU with a seeded stack."""
    f, h, w = out.shape
    p = _synth_params(cfg)
    if not lib().pt_synth_device(out.data_ptr(), int(seed), w, h, f, int(emitters), p.ctypes.data):
        raise RuntimeError("pt_synth_device failed")
    return out


def synth_emitters_device(out, seed, height, width, **cfg):
    """Ground-truth emitters into the torch float32 tensor `out` (frames, emitters, 5)."""
    f, e, _ = out.shape
    p = _synth_params(cfg)
    if not lib().pt_synth_emitters_device(out.data_ptr(), int(seed), width, height, f, e, p.ctypes.data):
        raise RuntimeError("pt_synth_emitters_device failed")
    return out


def localization_device(d_stack, d_out, threshold, watershed_ratio=0.0, fit=2, fit_params=(7, 1.3, 1.3, 0),
                        max_per_plane=None):
    """Localization over HBM-resident torch tensors; returns the number of rows written to d_out."""
    z, h, w = d_stack.shape
    cap = d_out.numel()
    mpp = (cap // NCOL) // max(z, 1) if max_per_plane is None else int(max_per_plane)
    p = _params(fit_params)
    n = lib().pt_localization_device(d_stack.data_ptr(), d_out.data_ptr(), cap, h, w, z, float(threshold),
                                     float(watershed_ratio), int(fit), p.ctypes.data, mpp)
    return n // NCOL


def crlb_device(d_rows, height, width, fit, fit_params, d_out, camera=(1.0, 0.0, 0.0)):
    """pt_crlb_device: bounds of X, Y, SigmaX, SigmaY, Theta, Amplitude, Offset, Intensity (n x 8 into the
    float64 tensor d_out) for the fitted rows in the float64 tensor d_rows (n x 18)."""
    n = d_rows.numel() // NCOL
    p = _params(fit_params)
    cam = np.asarray(camera, np.float64)
    if not lib().pt_crlb_device(d_rows.data_ptr(), n, height, width, int(fit), p.ctypes.data, cam.ctypes.data,
                                d_out.data_ptr()):
        raise RuntimeError("pt_crlb_device failed (see stderr)")
    return d_out


def batch_localization(stacks):
    """pt_batch_localization over a list of dicts {d_stack (planes, h, w) u16 tensor, d_out float64 tensor,
    threshold, watershed_ratio, fit, fit_params, max_per_plane}: every stack detected on its own, the fit
    jobs of all stacks fitted per fit class in one launch each. Returns the rows written per stack."""
    n = len(stacks)
    arr = (BatchStack * n)()
    keep = []
    for i, st in enumerate(stacks):
        d, o = st["d_stack"], st["d_out"]
        z, h, w = d.shape
        p = _params(st.get("fit_params", (7, 1.3, 1.3, 0)))
        keep.append(p)
        cap = o.numel()
        mpp = st.get("max_per_plane") or (cap // NCOL) // max(z, 1)
        arr[i] = BatchStack(d.data_ptr(), z, h, w, float(st["threshold"]), float(st.get("watershed_ratio", 0.0)),
                            int(st.get("fit", 2)), p.ctypes.data, o.data_ptr(), cap, int(mpp), 0)
    if not lib().pt_batch_localization(arr, n):
        raise RuntimeError("pt_batch_localization failed (see stderr)")
    return [arr[i].doublesWritten // NCOL for i in range(n)]


def batch_schedules():
    """Fit classes of the last batch: dicts {mode, roi, jobs, waves, refill, handover, ms, tuned}."""
    buf = (ctypes.c_double * 8 * 64)()
    k = lib().pt_batch_schedules(ctypes.cast(buf, _vp), 8 * 64)
    keys = ["mode", "roi", "jobs", "waves", "refill", "handover", "ms", "tuned"]
    out = []
    for i in range(min(k, 64)):
        row = [buf[i][j] for j in range(8)]
        d = dict(zip(keys, row))
        for key in ("mode", "roi", "jobs", "waves", "refill", "handover"):
            d[key] = int(d[key])
        d["tuned"] = bool(d["tuned"])
        out.append(d)
    return out


def fitting_noalloc(d_stack, d_positions, d_out, fit=2, fit_params=(7, 1.3, 1.3, 0)):
    """pt_fitting_noalloc: Fitting over HBM-resident torch tensors (stack (planes, h, w) u16, positions
    (n, 6) float64, output at least 18 * n float64) without any allocation; returns the number of rows."""
    z, h, w = d_stack.shape
    n = d_positions.numel() // 6
    p = _params(fit_params)
    return lib().pt_fitting_noalloc(d_stack.data_ptr(), d_positions.data_ptr(), n, d_out.data_ptr(), h, w, z, int(fit),
                                    p.ctypes.data) // NCOL


def astig_estimate(sigmas, pixel_size, model, z_max):
    """Astigmatism3DEstimation: (n, 2) observed (sigma_x, sigma_y) in pixels -> (n, 2) [z nm, Q_Z]."""
    s = np.ascontiguousarray(sigmas, np.float64).reshape(-1, 2)
    m = np.ascontiguousarray(model, np.float64).ravel()
    if m.size != 10:
        raise ValueError("the astigmatic model has 10 parameters (2 axes x [z0, w, c3, c4, A])")
    out = np.zeros_like(s)
    lib().Astigmatism3DEstimation(s.ctypes.data, out.ctypes.data, s.shape[0], float(pixel_size), m.ctypes.data,
                                  float(z_max))
    return out


def astig_estimate_device(d_sigmas, d_out, pixel_size, model, z_max):
    """Astigmatism3DEstimation over HBM-resident float64 torch tensors of shape (n, 2)."""
    m = np.ascontiguousarray(model, np.float64).ravel()
    if m.size != 10:
        raise ValueError("the astigmatic model has 10 parameters (2 axes x [z0, w, c3, c4, A])")
    n = d_sigmas.numel() // 2
    if not lib().pt_astig_estimate_device(d_sigmas.data_ptr(), d_out.data_ptr(), n, float(pixel_size), m.ctypes.data,
                                          float(z_max)):
        raise RuntimeError("pt_astig_estimate_device failed")
    return d_out


class Server:
    """Serving loop over device-resident frame chunks (pt_server_*): `slots` engines run requests
    side by side; requests are admitted in arrival order as slots and the HBM budget allow."""

    def __init__(self, slots=4, budget_bytes=32 << 30, chunk_pixels=0):
        self._h = lib().pt_server_create(int(slots), int(budget_bytes), int(chunk_pixels))
        if not self._h:
            raise RuntimeError("pt_server_create failed")

    def submit(self, d_stack, d_out, threshold, watershed_ratio=0.0, fit=2, fit_params=(7, 1.3, 1.3, 0),
               max_per_plane=None):
        """Queues Localization of d_stack (planes, h, w) u16 into d_out (float64); returns the id."""
        z, h, w = d_stack.shape
        cap = d_out.numel()
        mpp = (cap // NCOL) // max(z, 1) if max_per_plane is None else int(max_per_plane)
        p = _params(fit_params)
        rid = lib().pt_server_submit(self._h, d_stack.data_ptr(), z, h, w, float(threshold), float(watershed_ratio),
                                     int(fit), p.ctypes.data, p.size, d_out.data_ptr(), cap, mpp)
        if rid < 0:
            raise RuntimeError("pt_server_submit failed")
        return rid

    def wait(self, rid):
        """(rows, times) of a request; times = (submitted, started, done) ms on the server clock, slot."""
        n = ctypes.c_uint64(0)
        t = (ctypes.c_double * 4)()
        if not lib().pt_server_wait(self._h, int(rid), ctypes.byref(n), t):
            raise RuntimeError(f"request {rid} failed")
        return n.value // NCOL, (t[0], t[1], t[2], int(t[3]))

    def request_bytes(self, planes, height, width, out_doubles, max_per_plane):
        """HBM the admission counts for such a request (frames, rows and engine work space)."""
        return lib().pt_server_request_bytes(self._h, int(planes), int(height), int(width), int(out_doubles),
                                             int(max_per_plane))

    def now(self):
        return lib().pt_server_now(self._h)

    def close(self):
        if self._h:
            lib().pt_server_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
