#!/usr/bin/env python3
"""Fitted localizations/s of the detection + fitting path on 512x512 frames, one MI355X.

Workload (the 512^2 headline configuration): 10,000 seeded frames of 512x512 u16, 200 emitters per
frame (ADU = 100 + 2 * Poisson(20 + sum of pixel-integrated Gaussians, sigma 1.3 px, photons
log-normal median 2000, sigma_log 0.5)), a trous detection with the threshold AutoThreshold returns
on frame 0, watershed ratio 0, then the 7x7 Gaussian LM fit with free sigma (fit mode 2,
fit_params [7, 1.3, 1.3, 0], at most 20 iterations as the configuration states; the rate at the
reference's own cap of 75 iterations is reported beside it). The stack is generated on the device and
stays in HBM; one step is one Localization over the whole stack (detection, segmentation, fit, row
assembly), rows written to HBM. `value` counts fitted rows (rows whose fit did not fail).

Prints one JSON line. `roofline` is the dominant kernel (the detection kernel) measured with HIP
events on the library's stream, `roofline_fit` the fit stage (FP32 kernel and FP64 hand-over kernel);
`cpu_baseline` times the compiled reference (oracle/_ref, whose cap is 75 iterations) or, without it,
the CPU oracle, on a bounded sample of the same frames on this host's cores.
"""
import argparse
import ctypes
import json
import os
import sys
import time

# The serving leg runs four slot engines of three streams each; HIP maps streams onto this many
# hardware queues (4 by default), and streams that share a queue wait for each other.
os.environ.setdefault("GPU_MAX_HW_QUEUES", "16")

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "palmtracer-cpp_amd"))

HBM_PEAK_GBS = 8000.0
FP32_PEAK_TFLOPS = 157.3
FP64_PEAK_TFLOPS = 78.6
# Algorithmic work (per unit): a trous detection reads the u16 pixel, and for each mask pixel
# (thresholded potential above zero) writes its 4-byte potential and its 4-byte entry of the frame's
# mask list (the seed list and bitmap add a few bytes per local maximum, left out); the 7x7 /
# 5-parameter LM pass is 49 * 60 FLOP with the Jacobian and alpha/beta accumulation, 49 * 15 FLOP for
# a trial model evaluation.
ATROUS_READ_BYTES_PER_PX = 2.0
ATROUS_BYTES_PER_MASK_PX = 8.0
JAC_FLOP_7x7 = 2940.0
TRIAL_FLOP_7x7 = 735.0


def measured_traffic_per_px(size):
    """HBM bytes per pixel of the detection kernel from the committed rocprofv3 FETCH_SIZE /
    WRITE_SIZE summary (measured on 512x512 frames); None for other frame sizes."""
    path = os.path.join(ROOT, "profiles", "atrous_detect_kernel_512x512_traffic.md")
    if size != 512 or not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("traffic_bytes_per_pixel:"):
            return float(line.split(":")[1])
    return None


def cpu_baseline(stack_host, thr, fit_params, budget_s):
    """Reference (or oracle) Localization on the first frames of the stack, OpenMP over frames."""
    import numpy as np
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    R = oracle.ref_lib()
    kind = "reference" if R is not None else "port"
    fp = np.asarray(fit_params, np.float64)

    def run(nf):
        st = np.ascontiguousarray(stack_host[:nf])
        cap = nf * 1000 * 18
        out = np.empty(cap)
        t = time.perf_counter()
        if R is not None:
            n = R.ref_localization(st.ctypes.data, out.ctypes.data, cap, st.shape[1], st.shape[2], nf, thr, 0.0, 2,
                                   fp.ctypes.data)
        else:
            n = oracle.lib().or_localization(st.ctypes.data, out.ctypes.data, cap, st.shape[1], st.shape[2], nf, thr,
                                             0.0, 2, fp.ctypes.data, 75)
        return n // 18, time.perf_counter() - t

    nf = min(len(stack_host), 4 * threads)
    rows, dt = run(nf)
    per = dt / nf
    nf = int(min(len(stack_host), max(nf, budget_s / max(per, 1e-6))))
    rows, dt = run(nf)
    return {"value": rows / dt, "unit": "localizations/s", "cores": threads, "kind": kind,
            "sample": f"first {nf} frames of the same stack ({rows} rows in {dt:.2f} s), OpenMP over frames, "
                      f"LM cap 75 iterations (the reference's own)"}


def fit_precision(pt, torch, stack, thr, fit_params, frames):
    """The fast fit (the bench's) against the reference-order FP64 fit (equal to the CPU oracle within
    the fit tolerance in the parity tests) on the first frames of the stack."""
    sub = stack[:frames]
    out = torch.empty(frames * 1000 * 18, dtype=torch.float64, device="cuda")
    res = {}
    for prec in (0, 1):
        pt.set_option("fit_fp64", prec)
        n = pt.localization_device(sub, out, thr, 0.0, 2, fit_params, 1000)
        res[prec] = out[: n * 18].view(-1, 18).cpu().numpy().copy()
    pt.set_option("fit_fp64", 0)
    return dict(compare_fits(res[0], res[1]), frames=frames)


def serve_cameras(pt, torch, n_cams, sizes, load, rng, pool_frames=128, only=None):
    """The simulated cameras of the streamed load (seeded): frame side `sizes[c % len(sizes)]`, a frame
    rate uniform in [100, 1000] fps (times `load`), chunks of 10-100 frames, 0.5-3 emitters per um^2 at
    100 nm pixels capped at 5000 per frame, a pool of generated frames on the device. `only`: frame
    sides to build (the others are drawn but skipped)."""
    cams = []
    for c in range(n_cams):
        size = sizes[c % len(sizes)]
        fps = rng.uniform(100.0, 1000.0) * load
        chunk = rng.randint(10, 100)
        density = rng.uniform(0.5, 3.0)
        emitters = int(min(5000, round(density * size * size * 1e-2)))
        if only and size not in only:
            continue
        pool = torch.empty((max(pool_frames, chunk), size, size), dtype=torch.uint16, device="cuda")
        pt.synth_device(pool, seed=0x5EED1000 + c, emitters=emitters)
        thr = pt.autothreshold_device(pool[0])
        mpp = int(emitters * 1.25) + 64
        outs = [torch.empty(chunk * mpp * 18, dtype=torch.float64, device="cuda") for _ in range(8)]
        cams.append(dict(cam=c, size=size, fps=fps, chunk=chunk, density=density, emitters=emitters, pool=pool, thr=thr,
                         mpp=mpp, outs=outs, off=0, k=0, busy=[None] * len(outs)))
    torch.cuda.synchronize()
    return cams


def serve_alone(pt, torch, n_cams, sizes, fit_params, reps, chunk_pixels, only=None):
    """Each camera's chunk alone: one Localization over one chunk with the GPU otherwise idle, in the
    serving slots' configuration (chunk_pixels, flagged frames resolved in line); the sum over the
    cameras of (alone time x chunks per second) is the GPU time the offered load asks for."""
    import random
    rng = random.Random(0x5EED0005)
    cams = serve_cameras(pt, torch, n_cams, sizes, 1.0, rng, pool_frames=0, only=only)
    pt.set_option("chunk_pixels", chunk_pixels)
    pt.set_option("defer_flagged", 0)
    out = []
    try:
        for cam in cams:
            d = cam["pool"][: cam["chunk"]]
            o = cam["outs"][0]
            pt.localization_device(d, o, cam["thr"], 0.0, 2, fit_params, cam["mpp"])
            torch.cuda.synchronize()
            ts = []
            for _ in range(reps):
                t0 = time.perf_counter()
                n = pt.localization_device(d, o, cam["thr"], 0.0, 2, fit_params, cam["mpp"])
                torch.cuda.synchronize()
                ts.append(1e3 * (time.perf_counter() - t0))
            st = pt.last_stats()
            ms = min(ts)
            out.append({"camera": cam["cam"], "size": cam["size"], "frames": cam["chunk"], "emitters": cam["emitters"],
                        "rows": n, "ms": ms, "gpu_ms_per_s": ms * cam["fps"] / cam["chunk"],
                        "flagged_frames": st["flagged_frames"], "mask_pixels_per_frame": st["mask_pixels"] / cam["chunk"]})
            del cam["pool"], cam["outs"]
    finally:
        pt.set_option("chunk_pixels", 0)
        pt.set_option("defer_flagged", 2)
    return {"requests": out, "gpu_ms_per_s": sum(r["gpu_ms_per_s"] for r in out),
            "note": "each camera's chunk alone (min of the repetitions); gpu_ms_per_s = ms x the camera's chunks per second"}


def serve_latency(pt, torch, seconds, slots, n_cams, sizes, fit_params, load=1.0, chunk_pixels=1 << 27, budget_gb=192):
    """Streamed load: n_cams simulated cameras, each with a fixed frame size (`sizes` in turn), a frame
    rate uniform in [100, 1000] fps (times `load`), chunks of 10-100 frames and an emitter density
    uniform in 0.5-3 per um^2 at 100 nm pixels, capped at 5000 emitters per frame (all seeded). Chunks
    arrive as Poisson processes (mean rate = fps / chunk frames) and are submitted to a pt_server with
    `slots` slots; a chunk's latency is the time from its arrival to its rows in HBM, counted once per
    frame."""
    import random
    import threading
    import numpy as np
    rng = random.Random(0x5EED0005)
    cams = serve_cameras(pt, torch, n_cams, sizes, load, rng)
    budget = int(budget_gb) << 30
    srv = pt.Server(slots=slots, budget_bytes=budget, chunk_pixels=chunk_pixels)

    done = set()

    def take(cam):
        """The camera's next chunk of frames and a free output buffer (one whose last request is done)."""
        n = cam["chunk"]
        if cam["off"] + n > cam["pool"].shape[0]:
            cam["off"] = 0
        d = cam["pool"][cam["off"]: cam["off"] + n]
        cam["off"] += n
        b = cam["k"] % len(cam["outs"])
        cam["k"] += 1
        while cam["busy"][b] is not None and cam["busy"][b] not in done:
            time.sleep(0.0002)
        return d, cam["outs"][b], b

    def submit(cam):
        d, o, b = take(cam)
        rid = srv.submit(d, o, cam["thr"], 0.0, 2, fit_params, cam["mpp"])
        cam["busy"][b] = rid
        return rid

    # Warm-up: every slot sees the largest chunks, so that no buffer grows during the timed phase.
    warm = []
    for cam in sorted(cams, key=lambda c: -c["size"] * c["size"] * c["chunk"])[: 2 * slots]:
        warm.append(submit(cam))
    for r in warm:
        srv.wait(r)
        done.add(r)

    arrivals = []
    for ci, cam in enumerate(cams):
        t, lam = 0.0, cam["fps"] / cam["chunk"]
        while True:
            t += rng.expovariate(lam)
            if t >= seconds:
                break
            arrivals.append((t, ci))
    arrivals.sort()
    submitted = []
    base = srv.now() + 50.0
    t_base = time.perf_counter() + 0.05

    def scheduler():
        for t, ci in arrivals:
            delay = t_base + t - time.perf_counter()
            if delay > 0:
                time.sleep(delay)
            rid = submit(cams[ci])
            submitted.append((rid, base + 1e3 * t, cams[ci]["chunk"]))

    th = threading.Thread(target=scheduler)
    th.start()
    lat, weights, rows, lags, i, sizes_of = [], [], 0, [], 0, []
    while i < len(arrivals):
        if i >= len(submitted):
            time.sleep(0.0002)
            continue
        rid, planned, n = submitted[i]
        r, (t_sub, t_start, t_done, _) = srv.wait(rid)
        done.add(rid)
        lat.append(t_done - planned)
        lags.append(t_sub - planned)
        weights.append(n)
        sizes_of.append(cams[arrivals[i][1]]["size"])
        rows += r
        i += 1
    th.join()
    t_end = srv.now()
    srv.close()
    lat, weights, sizes_of = np.array(lat), np.array(weights, dtype=np.float64), np.array(sizes_of)

    def pct(sel, q):
        l, w = lat[sel], weights[sel]
        order = np.argsort(l)
        cw = np.cumsum(w[order]) / w.sum()
        return float(l[order][min(int(np.searchsorted(cw, q)), len(l) - 1)])

    allsel = np.ones(len(lat), bool)
    frames = int(weights.sum())
    px = sum(cams[ci]["chunk"] * cams[ci]["size"] ** 2 for _, ci in arrivals)
    by_size = {}
    for sz in sorted(set(sizes_of.tolist())):
        sel = sizes_of == sz
        by_size[f"{sz}x{sz}"] = {"p50_ms": pct(sel, 0.5), "p99_ms": pct(sel, 0.99), "frames": int(weights[sel].sum())}
    # Fast fit against the reference-order FP64 fit on one chunk of every camera size.
    precision = {}
    for cam in {c["size"]: c for c in cams}.values():
        d = cam["pool"][:10]
        o = torch.empty(10 * cam["mpp"] * 18, dtype=torch.float64, device="cuda")
        res = {}
        for prec in (0, 1):
            pt.set_option("fit_fp64", prec)
            k = pt.localization_device(d, o, cam["thr"], 0.0, 2, fit_params, cam["mpp"])
            res[prec] = o[: k * 18].view(-1, 18).cpu().numpy().copy()
        pt.set_option("fit_fp64", 0)
        precision[f"{cam['size']}x{cam['size']}"] = compare_fits(res[0], res[1])
    return {
        "p50_ms": pct(allsel, 0.50), "p99_ms": pct(allsel, 0.99), "max_ms": float(lat.max()),
        "latency": "chunk arrival to its rows in HBM, per frame", "by_frame_size": by_size,
        "requests": len(arrivals), "frames": frames, "rows": rows,
        "offered_frames_per_s": sum(c["fps"] for c in cams), "offered_gpx_per_s": px / seconds / 1e9,
        "localizations_per_s": rows / ((t_end - base) * 1e-3), "load": load,
        "cameras": [{"size": c["size"], "fps": round(c["fps"], 1), "chunk": c["chunk"], "density_per_um2": round(c["density"], 2),
                     "emitters": c["emitters"]} for c in cams],
        "slots": slots, "budget_gb": budget >> 30, "chunk_pixels": chunk_pixels,
        "seconds": seconds, "max_submit_lag_ms": float(max(lags)),
        "fit_precision_fast_vs_fp64_10_frames": precision,
    }


def astig3d_leg(pt, torch, frames, steps):
    """Astigmatic 3D workload: `frames` sCMOS-like 2048^2 frames (per-pixel offset ~ N(100, 3) ADU, gain
    ~ N(2, 0.1)), 2000 emitters per frame with z ~ U[-600, 600] nm and sigma_{x,y} = 1.2 sqrt(1 +
    ((z -/+ 250) / 400)^2) px; one step = Localization with the 9 x 9 fit of sigma_x / sigma_y (fit
    mode 3, fast fit) and Astigmatism3DEstimation of every row's widths, both on the device."""
    import numpy as np
    cfg = dict(astig=1, s0=1.2, c=250.0, d=400.0, zmin=-600.0, zmax=600.0, scmos=1, sd_offset=3.0, sd_gain=0.1)
    fit_params = (9.0, 1.4, 1.4, 0.0)
    model = [250.0, 400.0, 0.0, 0.0, 120.0, -250.0, 400.0, 0.0, 0.0, 120.0]
    d = torch.empty((frames, 2048, 2048), dtype=torch.uint16, device="cuda")
    pt.synth_device(d, seed=0x5EED0003, emitters=2000, **cfg)
    thr = pt.autothreshold_device(d[0], (9.0,))
    mpp = 4000
    out = torch.empty(frames * mpp * 18, dtype=torch.float64, device="cuda")
    zs = torch.empty((frames * mpp, 2), dtype=torch.float64, device="cuda")

    def step():
        n = pt.localization_device(d, out, thr, 0.0, 3, fit_params, mpp)
        rows = out[: n * 18].view(-1, 18)
        sig = rows[:, 8:10].contiguous()
        pt.astig_estimate_device(sig, zs[:n], 100.0, model, 600.0)
        return n

    step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        n = step()
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / steps
    st = pt.stage_times()
    handed = pt.last_stats()["fit_handed_over"]
    rows = out[: n * 18].view(-1, 18)
    fitted = int((rows[:, 13] > 0).sum().item())
    # Fast fit against the reference-order FP64 fit on the first 20 frames of the stack.
    sub = d[:20]
    res = {}
    o2 = torch.empty(20 * mpp * 18, dtype=torch.float64, device="cuda")
    for prec in (0, 1):
        pt.set_option("fit_fp64", prec)
        k = pt.localization_device(sub, o2, thr, 0.0, 3, fit_params, mpp)
        res[prec] = o2[: k * 18].view(-1, 18).cpu().numpy().copy()
    pt.set_option("fit_fp64", 0)
    prec = compare_fits(res[0], res[1], sigma_y=True)
    # The precision extension on the same frames (the first 100): the Poisson likelihood fit (camera
    # model 2 ADU per photon, 100 ADU offset, read variance 9 ADU^2, the spread of the sCMOS offsets),
    # the FP64 Cramer-Rao bounds of every fitted row, and z from the likelihood fit's widths.
    cam = (2.0, 100.0, 9.0)
    nl = min(frames, 100)
    sub = d[:nl]
    o3 = torch.empty(nl * mpp * 18, dtype=torch.float64, device="cuda")
    bounds = torch.empty((nl * mpp, 8), dtype=torch.float64, device="cuda")
    z3 = torch.empty((nl * mpp, 2), dtype=torch.float64, device="cuda")
    pt.set_option("fit_objective", 1)
    pt.set_option("camera_gain", cam[0])
    pt.set_option("camera_offset", cam[1])
    pt.set_option("camera_read_var", cam[2])
    try:
        def mle_step():
            k = pt.localization_device(sub, o3, thr, 0.0, 3, fit_params, mpp)
            rows3 = o3[: k * 18].view(-1, 18)
            pt.crlb_device(rows3, 2048, 2048, 3, fit_params, bounds[:k], cam)
            pt.astig_estimate_device(rows3[:, 8:10].contiguous(), z3[:k], 100.0, model, 600.0)
            return k
        mle_step()
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        for _ in range(steps):
            k3 = mle_step()
        torch.cuda.synchronize()
        dt3 = (time.perf_counter() - t1) / steps
    finally:
        pt.set_option("fit_objective", 0)
        pt.set_option("camera_gain", 1.0)
        pt.set_option("camera_offset", 0.0)
        pt.set_option("camera_read_var", 0.0)
    rows3 = o3[: k3 * 18].view(-1, 18)
    ok3 = (rows3[:, 13] > 0) & torch.isfinite(bounds[:k3, 0])
    fitted3 = int(ok3.sum().item())
    b = bounds[:k3][ok3]
    mle = {"frames": nl, "ms_per_step": dt3 * 1e3, "localizations_per_s": fitted3 / dt3, "fitted_rows": fitted3,
           "camera": {"gain_adu_per_photon": cam[0], "offset_adu": cam[1], "read_variance_adu2": cam[2]},
           "median_crlb": {"x_px": float(b[:, 0].median()), "y_px": float(b[:, 1].median()),
                           "sigma_x_px": float(b[:, 2].median()), "intensity_rel": float((b[:, 7] / rows3[ok3, 7]).median())},
           "step": "likelihood fit (fit_objective 1, FP64, one ROI per lane) + pt_crlb_device + Astigmatism3DEstimation"}
    return {"workload": f"{frames} frames 2048x2048 u16 sCMOS (offset N(100,3), gain N(2,0.1)), 2000 astigmatic "
                        f"emitters/frame (z U[-600,600] nm, s0 1.2 px, c 250 nm, d 400 nm), 9x9 fit mode 3 + "
                        f"Astigmatism3DEstimation", "frames": frames, "steps": steps, "ms_per_step": dt * 1e3,
            "localizations_per_s": fitted / dt, "rows": n, "fitted_rows": fitted, "threshold": thr,
            "stage_ms": {"atrous": st["atrous_ms"], "segment": st["segment_ms"], "fit": st["fit_ms"],
                         "assemble": st["assemble_ms"]},
            "fit_handed_over": handed, "fit_precision_first_20_frames": prec, "likelihood_and_bounds": mle}


def batch_leg(pt, torch, frame_scale, steps, per_frame_sample):
    """Batched uneven workload: 64 seeded stacks of frame sizes drawn from {128, 256, 512, 1024, 2048}^2,
    frame counts uniform in 50-2000 (times `frame_scale`), emitter density uniform in 0.1-5 per um^2 at
    100 nm pixels, capped at 5000 emitters per frame. Each stack gets the 7x7 fit of sigma (mode 2) or
    the 9x9 astigmatic fit of sigma_x / sigma_y (mode 3), assigned so that the two classes hold about
    equal numbers of emitters. One step = pt_batch_localization over all stacks (per-stack detection,
    one fit launch per class in a captured graph, schedules from the autotuner). The same stacks are
    then run through the plain per-stack path (one Localization per stack) and the plain per-frame path
    (one Localization per frame, on the first `per_frame_sample` frames of every stack)."""
    import numpy as np
    rng = np.random.default_rng(0x5EED0004)
    astig_cfg = dict(astig=1, s0=1.2, c=250.0, d=400.0, zmin=-600.0, zmax=600.0)
    p2d, p3d = (7.0, 1.3, 1.3, 0.0), (9.0, 1.4, 1.4, 0.0)
    load = [0.0, 0.0]
    stacks = []
    for k in range(64):
        side = int(rng.choice([128, 256, 512, 1024, 2048]))
        frames = max(1, int(round(int(rng.integers(50, 2001)) * frame_scale)))
        density = float(rng.uniform(0.1, 5.0))
        em = int(min(5000, round(density * side * side * 1e-2)))
        cls = 0 if load[0] <= load[1] else 1
        load[cls] += frames * em
        d = torch.empty((frames, side, side), dtype=torch.uint16, device="cuda")
        pt.synth_device(d, seed=0x5EED4000 + k, emitters=em, **(astig_cfg if cls else {}))
        fp = p3d if cls else p2d
        thr = pt.autothreshold_device(d[0], (fp[0],))
        mpp = int(em * 1.25) + 64
        out = torch.empty(frames * mpp * 18, dtype=torch.float64, device="cuda")
        stacks.append(dict(d_stack=d, d_out=out, threshold=thr, fit=3 if cls else 2, fit_params=fp, max_per_plane=mpp,
                           side=side, frames=frames, emitters=em, density=density))
    torch.cuda.synchronize()

    def fitted_of(rows):
        n = 0
        for st, r in zip(stacks, rows):
            n += int((st["d_out"][: r * 18].view(-1, 18)[:, 13] > 0).sum().item())
        return n

    pt.batch_localization(stacks)   # warm-up: first sight of each class runs the autotuner
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        rows = pt.batch_localization(stacks)
        st_times = pt.stage_times()
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / steps
    fitted = fitted_of(rows)
    sched = pt.batch_schedules()
    batch_fit_ms = st_times["fit_ms"]

    # Plain per-stack path: one Localization per stack (same fast fit, default schedule).
    plain_rows = []
    t1 = time.perf_counter()
    fit_ms_stacks = 0.0
    for st in stacks:
        plain_rows.append(pt.localization_device(st["d_stack"], st["d_out"], st["threshold"], 0.0, st["fit"],
                                                 st["fit_params"], st["max_per_plane"]))
        fit_ms_stacks += pt.stage_times()["fit_ms"]
    torch.cuda.synchronize()
    dt_stack = time.perf_counter() - t1
    fitted_stack = fitted_of(plain_rows)

    # Plain per-frame path: one Localization per frame, the first frames of every stack.
    n_frames, fitted_frame = 0, 0
    sub = []
    for st in stacks:
        m = min(per_frame_sample, st["frames"])
        o = torch.empty(st["max_per_plane"] * 18, dtype=torch.float64, device="cuda")
        sub.append((st, m, o))
    for st, m, o in sub:   # warm every frame size once
        pt.localization_device(st["d_stack"][:1], o, st["threshold"], 0.0, st["fit"], st["fit_params"], st["max_per_plane"])
    torch.cuda.synchronize()
    t2 = time.perf_counter()
    counts = []
    for st, m, o in sub:
        for f in range(m):
            r = pt.localization_device(st["d_stack"][f: f + 1], o, st["threshold"], 0.0, st["fit"], st["fit_params"],
                                       st["max_per_plane"])
            counts.append(int((o[: r * 18].view(-1, 18)[:, 13] > 0).sum().item()) if r else 0)
            n_frames += 1
    torch.cuda.synchronize()
    dt_frame = time.perf_counter() - t2
    # The sync of each count is part of a per-frame caller's loop (rows come back per call).
    fitted_frame = sum(counts)
    rate = fitted / dt
    rate_stack = fitted_stack / dt_stack
    rate_frame = fitted_frame / dt_frame
    total_frames = sum(st["frames"] for st in stacks)
    gpx = sum(st["frames"] * st["side"] ** 2 for st in stacks) / 1e9
    return {"workload": "64 stacks, frame sides from {128,256,512,1024,2048}, frame counts U[50,2000] x "
                        f"{frame_scale:g}, 0.1-5 emitters/um^2 at 100 nm px capped 5000/frame, 7x7 mode 2 / 9x9 "
                        "astigmatic mode 3 (fast fit), seed 0x5EED0004",
            "frame_scale": frame_scale, "stacks": len(stacks), "frames": total_frames, "gpx": gpx,
            "emitters_by_class": {"7x7_mode2": load[0], "9x9_mode3": load[1]},
            "steps": steps, "ms_per_step": dt * 1e3, "rows": int(sum(rows)), "fitted_rows": fitted,
            "localizations_per_s": rate, "fit_stage_ms": batch_fit_ms,
            "schedules": sched,
            "per_stack": {"ms": dt_stack * 1e3, "fitted_rows": fitted_stack, "localizations_per_s": rate_stack,
                          "fit_stage_ms_sum": fit_ms_stacks},
            "per_frame": {"frames": n_frames, "frames_per_stack": per_frame_sample, "ms": dt_frame * 1e3,
                          "fitted_rows": fitted_frame, "localizations_per_s": rate_frame},
            "speedup_vs_per_frame": rate / rate_frame, "speedup_vs_per_stack": rate / rate_stack,
            "stack_sides": [st["side"] for st in stacks], "stack_frames": [st["frames"] for st in stacks]}


def host_path_leg(pt, torch, stack, thr, fit_params, max_per_plane, steps):
    """The reference's own entry point over host buffers: Localization (CPU/main.cpp:112-121) of the
    whole stack from pageable host memory into a host output, streamed through the device in pieces
    (pinned staging, uploads and downloads overlapping the computation). End to end, PCIe included."""
    import numpy as np
    F, H, W = stack.shape
    host = stack.cpu().numpy()
    out = np.empty(F * max_per_plane * 18, np.float64)
    params = np.asarray(fit_params, np.float64)
    L = pt.lib()

    def call():
        return L.Localization(host.ctypes.data, out.ctypes.data, out.size, H, W, F, thr, 0.0, 2, params.ctypes.data)

    call()
    t0 = time.perf_counter()
    for _ in range(steps):
        n = call()
    dt = (time.perf_counter() - t0) / steps
    st = pt.stage_times()
    rows = out[:n].reshape(-1, 18)
    fitted = int((rows[:, 13] > 0).sum())
    gb = F * H * W * 2 / 1e9
    return {"entry": "Localization (host buffers)", "frames": F, "steps": steps, "ms_per_call": dt * 1e3,
            "localizations_per_s": fitted / dt, "fitted_rows": fitted, "input_gb": gb, "input_gb_per_s": gb / dt,
            "h2d_ms": st["h2d_ms"], "d2h_ms": st["d2h_ms"], "output_gb": n * 8 / 1e9, "pieces": int(st["stream_pieces"]),
            "note": "h2d/d2h: device time of the copies of the last call, overlapped with the computation"}


def compare_fits(a, e, sigma_y=False):
    """Fast-fit rows `a` against FP64 rows `e` of the same frames: rows paired within each plane by
    position, largest errors over rows fitted on both sides."""
    import numpy as np
    X, Y, INT, SX, SY, AMP, OFF = 4, 5, 7, 8, 9, 13, 14
    pairs = []
    for p in np.unique(e[:, 1]):
        A, E = a[a[:, 1] == p], e[e[:, 1] == p]
        d = ((A[:, None, X:Y + 1] - E[None, :, X:Y + 1]) ** 2).sum(-1)
        pairs.append((A, E[d.argmin(1)]))
    A = np.concatenate([q[0] for q in pairs])
    E = np.concatenate([q[1] for q in pairs])
    both = (A[:, AMP] > 0) & (E[:, AMP] > 0)
    one = int(((A[:, AMP] > 0) != (E[:, AMP] > 0)).sum())
    A, E = A[both], E[both]
    rel = lambda c: np.abs(A[:, c] - E[:, c]) / np.maximum(np.abs(E[:, c]), 1e-300)
    cols = [SX, SY] if sigma_y else [SX]
    within = ((np.abs(A[:, X] - E[:, X]) <= 1e-3) & (np.abs(A[:, Y] - E[:, Y]) <= 1e-3) & (rel(INT) <= 1e-4) &
              (rel(AMP) <= 1e-4) & (rel(OFF) <= 1e-4))
    for c in cols:
        within &= rel(c) <= 1e-4
    return {"rows": int(len(e)), "rows_fast": int(len(a)), "fitted_on_one_side_only": one,
            "max_abs_dx_px": float(np.abs(A[:, X] - E[:, X]).max()), "max_abs_dy_px": float(np.abs(A[:, Y] - E[:, Y]).max()),
            "max_rel_intensity": float(rel(INT).max()), "max_rel_amplitude": float(rel(AMP).max()),
            "max_rel_sigma": float(max(rel(c).max() for c in cols)), "max_rel_offset": float(rel(OFF).max()),
            "fraction_within_tolerance": float(within.mean()),
            "tolerance": "x, y 1e-3 px; intensity, amplitude, sigma, offset 1e-4 relative"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--frames", type=int, default=10000)
    ap.add_argument("--size", type=int, default=512)
    ap.add_argument("--emitters", type=int, default=200)
    ap.add_argument("--precision", choices=["fp32", "fp64"], default="fp32",
                    help="fp32: the fast fit (FP32 kernel, FP64 hand-over); fp64: the reference-order FP64 fit")
    ap.add_argument("--max-iter", type=int, default=20, help="LM iteration cap (the configuration's 20)")
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--serve-seconds", type=float, default=4.0,
                    help="length of the streamed-load phase (0: skip it)")
    ap.add_argument("--serve-slots", type=int, default=4)
    ap.add_argument("--serve-cameras", type=int, default=16)
    ap.add_argument("--astig-frames", type=int, default=1000,
                    help="frames of the astigmatic 3D leg (2048^2 sCMOS; 0: skip it)")
    ap.add_argument("--astig-steps", type=int, default=3)
    ap.add_argument("--host-steps", type=int, default=2,
                    help="calls of the host-buffer Localization over the whole stack (0 skips the leg)")
    ap.add_argument("--batch-frame-scale", type=float, default=0.1,
                    help="scale of the batched workload's frame counts (U[50,2000] per stack); 0 skips the leg")
    ap.add_argument("--batch-steps", type=int, default=2)
    ap.add_argument("--batch-frame-sample", type=int, default=4,
                    help="frames per stack timed through the plain per-frame path")
    ap.add_argument("--serve-chunk-pixels", type=int, default=1 << 27,
                    help="pixels per engine chunk of a serving slot")
    ap.add_argument("--serve-budget-gb", type=int, default=192, help="HBM budget of the serving loop")
    ap.add_argument("--serve-alone", type=int, default=0, metavar="REPS",
                    help="only time each serving camera's chunk alone (REPS repetitions) and print that, then exit")
    ap.add_argument("--serve-alone-sizes", default="",
                    help="with --serve-alone: frame sides to time (default: every camera)")
    ap.add_argument("--serve-sizes", default="256,512,1024,2048",
                    help="frame sides of the cameras, assigned in turn")
    ap.add_argument("--serve-load", type=float, default=1.0,
                    help="factor on every camera's frame rate (1: rates uniform in 100-1000 fps)")
    ap.add_argument("--set", action="append", default=[], metavar="NAME=VALUE",
                    help="library option (pt_set_option) applied before the run, e.g. detect_rows=103")
    ap.add_argument("--copy-bytes", type=int, default=0,
                    help="only run one device-to-device copy of this many bytes (calibrates the HBM byte "
                         "counters of rocprofv3 --pmc against a transfer of known size) and exit")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: the library has no multi-GPU path (asked for --gpus {args.gpus}).",
              file=sys.stderr)
        sys.exit(2)

    import torch
    if args.copy_bytes > 0:
        src = torch.empty(args.copy_bytes, dtype=torch.uint8, device="cuda").fill_(1)
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes": args.copy_bytes}))
        return
    import palmtracer_amd as pt

    if args.serve_alone > 0:
        pt.set_option("fit_fp64", 0)
        pt.set_option("max_iter", args.max_iter)
        for kv in args.set:
            name, value = kv.split("=", 1)
            pt.set_option(name, float(value))
        sizes = [int(v) for v in args.serve_sizes.split(",")]
        only = [int(v) for v in args.serve_alone_sizes.split(",")] if args.serve_alone_sizes else None
        print(json.dumps({"serve_alone": serve_alone(pt, torch, args.serve_cameras, sizes, (7.0, 1.3, 1.3, 0.0),
                                                     args.serve_alone, args.serve_chunk_pixels, only)}))
        return

    F, S = args.frames, args.size
    fit_params = (7.0, 1.3, 1.3, 0.0)
    seed = 0x5EED0002
    stack = torch.empty((F, S, S), dtype=torch.uint16, device="cuda")
    pt.synth_device(stack, seed=seed, emitters=args.emitters)
    torch.cuda.synchronize()
    thr = pt.autothreshold_device(stack[0])
    max_per_plane = 1000
    out = torch.empty(F * max_per_plane * 18, dtype=torch.float64, device="cuda")
    pt.set_option("fit_fp64", 1 if args.precision == "fp64" else 0)
    pt.set_option("max_iter", args.max_iter)
    pt.set_option("profile", 1)
    options = {}
    for kv in args.set:
        name, value = kv.split("=", 1)
        pt.set_option(name, float(value))
        options[name] = float(value)

    for _ in range(args.warmup):
        pt.localization_device(stack, out, thr, 0.0, 2, fit_params, max_per_plane)
    torch.cuda.synchronize()
    stages = {"atrous_ms": 0.0, "segment_ms": 0.0, "fit_ms": 0.0, "assemble_ms": 0.0, "chunks": 0.0,
              "jacobian_passes": 0.0, "trial_passes": 0.0, "deferred_ms": 0.0, "group_jacobian_passes": 0.0,
              "group_trial_passes": 0.0}
    rows = 0
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        rows = pt.localization_device(stack, out, thr, 0.0, 2, fit_params, max_per_plane)
        st = pt.stage_times()
        for k in stages:
            stages[k] += st[k]
    last = pt.last_stats()
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    ms_step = 1e3 * dt / args.steps
    res = out[: rows * 18].view(-1, 18)
    fitted = int((res[:, 13] > 0).sum().item())

    # Roofline of the dominant stage's kernel (averages over all launches of the timed steps).
    launches = max(stages["chunks"], 1.0)
    atrous_ms = stages["atrous_ms"] / launches
    fit_ms = stages["fit_ms"] / launches
    px_per_launch = F * S * S / (stages["chunks"] / args.steps)
    if rows == 0:
        raise SystemExit("no localizations: the pipeline reported an error (see stderr)")
    # The fit stage of a call: the FP32 kernel over all ROIs of the call, then the FP64 group kernel
    # over the ROIs it handed over (the deferred frames' fits on the side lane add their passes, under
    # 1 % of them, to the counters but not their time). FLOPs on the reference's pass schedule.
    fit_ms = stages["fit_ms"] / args.steps
    flops32 = (stages["jacobian_passes"] * JAC_FLOP_7x7 + stages["trial_passes"] * TRIAL_FLOP_7x7) / args.steps
    flops64 = (stages["group_jacobian_passes"] * JAC_FLOP_7x7 + stages["group_trial_passes"] * TRIAL_FLOP_7x7) / args.steps
    if args.precision == "fp64":
        fit_kernel, fit_peak = "fit_kernel (FP64 LM in the reference's operation order, one ROI per lane)", FP64_PEAK_TFLOPS
    else:
        fit_kernel = ("fit_persistent_kernel (FP32 LM, one ROI per lane) + fit_group_kernel (FP64 LM, eight lanes "
                      "per ROI, the ROIs handed over)")
        fit_peak = FP32_PEAK_TFLOPS
    fit_ach = (flops32 + flops64) / (fit_ms * 1e-3) / 1e12
    roof_fit = {"bound": "valu", "kernel": fit_kernel,
                "achieved": fit_ach, "peak": fit_peak, "unit": "TFLOP/s", "frac": fit_ach / fit_peak,
                "launch_ms": fit_ms, "flop_per_launch": flops32 + flops64, "flop_fp64_part": flops64,
                "jacobian_passes": stages["jacobian_passes"] / args.steps,
                "trial_passes": stages["trial_passes"] / args.steps,
                "group_jacobian_passes": stages["group_jacobian_passes"] / args.steps,
                "group_trial_passes": stages["group_trial_passes"] / args.steps,
                "rois_handed_to_fp64": last["fit_handed_over"]}
    # The dominant kernel by device time: the detection kernel (HBM-bound).
    mask_per_launch = last["mask_pixels"] / max(stages["chunks"] / args.steps, 1.0)
    bytes_l = ATROUS_READ_BYTES_PER_PX * px_per_launch + ATROUS_BYTES_PER_MASK_PX * mask_per_launch
    ach = bytes_l / (atrous_ms * 1e-3) / 1e9
    tpp = measured_traffic_per_px(S)
    roof = {"bound": "hbm", "kernel": "atrous_detect_kernel", "achieved": ach, "peak": HBM_PEAK_GBS,
            "unit": "GB/s", "frac": ach / HBM_PEAK_GBS,
            "traffic": None if tpp is None else tpp * px_per_launch, "launch_ms": atrous_ms,
            "bytes_per_launch": bytes_l, "mask_pixels_per_launch": mask_per_launch}
    stage_ms = {k: v / args.steps for k, v in stages.items() if k.endswith("_ms")}

    # The detection launches alone: one main lane, so that no other lane's kernels share the GPU with
    # them (the timed steps overlap two lanes, and their events include that sharing).
    pt.set_option("lanes", 1)
    pt.localization_device(stack, out, thr, 0.0, 2, fit_params, max_per_plane)
    alone_ms, alone_launches = 0.0, 0.0
    for _ in range(2):
        pt.localization_device(stack, out, thr, 0.0, 2, fit_params, max_per_plane)
        st = pt.stage_times()
        alone_ms += st["atrous_ms"]
        alone_launches += st["chunks"]
    pt.set_option("lanes", options.get("lanes", 2.0))
    torch.cuda.synchronize()
    alone = alone_ms / max(alone_launches, 1.0)
    roof["standalone_launch_ms"] = alone
    roof["standalone_frac"] = bytes_l / (alone * 1e-3) / 1e9 / HBM_PEAK_GBS

    # The same workload at the reference's own iteration cap (75), fitted rows per second.
    alt_iter = 75 if args.max_iter != 75 else 20
    pt.set_option("max_iter", alt_iter)
    pt.localization_device(stack, out, thr, 0.0, 2, fit_params, max_per_plane)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    for _ in range(args.steps):
        n_alt = pt.localization_device(stack, out, thr, 0.0, 2, fit_params, max_per_plane)
    torch.cuda.synchronize()
    ms_alt = 1e3 * (time.perf_counter() - t1) / args.steps
    fitted_alt = int((out[: n_alt * 18].view(-1, 18)[:, 13] > 0).sum().item())
    other_cap = {"max_iter": alt_iter, "ms_per_step": ms_alt, "localizations_per_s": fitted_alt / (ms_alt * 1e-3)}

    # The library's default fit on the same workload: FP64 in the reference's operation order (rows
    # bit-equal to the reference's) at the reference's own cap, the drop-in Localization's rate.
    pt.set_option("fit_fp64", 1)
    pt.set_option("max_iter", 75)
    pt.localization_device(stack, out, thr, 0.0, 2, fit_params, max_per_plane)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    ref_steps = max(1, min(args.steps, 2))
    for _ in range(ref_steps):
        n_ref = pt.localization_device(stack, out, thr, 0.0, 2, fit_params, max_per_plane)
    torch.cuda.synchronize()
    ms_ref = 1e3 * (time.perf_counter() - t1) / ref_steps
    fitted_ref = int((out[: n_ref * 18].view(-1, 18)[:, 13] > 0).sum().item())
    library_default = {"fit": "FP64, the reference's operation order (library default)", "max_iter": 75,
                       "steps": ref_steps, "ms_per_step": ms_ref, "localizations_per_s": fitted_ref / (ms_ref * 1e-3),
                       "fitted_rows": fitted_ref}
    pt.set_option("fit_fp64", 0 if args.precision == "fp32" else 1)
    pt.set_option("max_iter", args.max_iter)

    precision = fit_precision(pt, torch, stack, thr, fit_params, min(F, 200)) if args.precision == "fp32" else None

    host_path = host_path_leg(pt, torch, stack, thr, fit_params, max_per_plane, args.host_steps) if args.host_steps > 0 else None

    serving = None
    if args.serve_seconds > 0:
        sizes = [int(v) for v in args.serve_sizes.split(",")]
        serving = serve_latency(pt, torch, args.serve_seconds, args.serve_slots, args.serve_cameras, sizes, fit_params,
                                args.serve_load, args.serve_chunk_pixels, args.serve_budget_gb)

    astig = astig3d_leg(pt, torch, args.astig_frames, args.astig_steps) if args.astig_frames > 0 else None

    batched = None
    if args.batch_frame_scale > 0:
        batched = batch_leg(pt, torch, args.batch_frame_scale, args.batch_steps, args.batch_frame_sample)

    cpu = None
    if not args.no_cpu_baseline:
        host = stack[: min(F, 4000)].cpu().numpy()
        cpu = cpu_baseline(host, thr, fit_params, args.cpu_seconds)

    line = {
        "metric": "fitted localizations/sec at 512^2 frames",
        "value": fitted / (ms_step * 1e-3),
        "unit": "localizations/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32" if args.precision == "fp32" else "f64",
        "data": "synthetic: seeded on-device generator (Philox), stack resident in HBM",
        "config": {"workload": f"{F} frames {S}x{S} u16, {args.emitters} emitters/frame, a trous + watershed "
                               f"(ratio 0) + 7x7 Gaussian LM fit mode 2 (max {args.max_iter} it), threshold "
                               f"{thr:.6f} (AutoThreshold of frame 0)", "max_iter": args.max_iter,
                   "other_iteration_cap": other_cap, "library_default_fit": library_default,
                   "frames": F, "height": S, "width": S, "emitters_per_frame": args.emitters, "fit": 2,
                   "roi": 7, "threshold": thr, "rows": rows, "fitted_rows": fitted, "seed": seed,
                   "stage_ms_per_step": stage_ms, "options": options,
                   "watershed_frames": {k: last[k] for k in ("flagged_frames", "deferred_frames", "flag_unproven",
                                                             "flag_overflow", "rank_too_big", "rank_equal_keys",
                                                             "rank_unsettled", "rank_two_orders", "heap_frames",
                                                             "multi_components")}},
        "roofline": roof,
        "roofline_fit": roof_fit,
        "fit_precision": precision,
        "host_path": host_path,
        "astig3d": astig,
        "batched": batched,
        "serving": serving,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
