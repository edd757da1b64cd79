"""Writes the committed fixtures of the parity tests (run in a container that has the reference tree).

    python tests/golden/make_golden.py

Outputs (compressed .npz, data only):
  reference_inputs.npz  stack.tif and stack_quadrant.tif of the reference's tests (Tests_CPU/Input),
                        decoded to (planes, height, width) u16.
  reference_csv.npz     the reference's golden localization CSVs (Tests_CPU/Input/ref/*.csv) as arrays.
  ref_vectors.npz       outputs of the compiled reference (oracle/_ref/libref.so, built by oracle/Makefile
                        from the reference sources) on seeded synthetic frames: label maps, region lists,
                        Localization rows for every fit mode with watershed on and off, Fitting rows and
                        AutoThreshold values. They pin the parts of the oracle that the reference's own
                        tests leave open (label maps on touching spots, fusion outcomes, LM trajectories).
"""
import glob
import os
import sys

import numpy as np
from PIL import Image

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
REF_INPUT = "/root/reference/Tests_CPU/Input"
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, HERE)

import oracle  # noqa: E402
import synth_np  # noqa: E402


def load_tiff(path):
    im = Image.open(path)
    frames = []
    for i in range(im.n_frames):
        im.seek(i)
        frames.append(np.array(im, dtype=np.uint16))
    return np.stack(frames)


def main():
    oracle.build()
    R = oracle.ref_lib()
    if R is None:
        raise SystemExit("oracle/_ref/libref.so is missing: the reference tree is needed to write the fixtures")

    stack = load_tiff(os.path.join(REF_INPUT, "stack.tif"))
    quad = load_tiff(os.path.join(REF_INPUT, "stack_quadrant.tif"))
    np.savez_compressed(os.path.join(HERE, "reference_inputs.npz"), stack=stack, stack_quadrant=quad)

    csv = {}
    for p in sorted(glob.glob(os.path.join(REF_INPUT, "ref", "*localizations*.csv"))):
        name = os.path.basename(p)[:-4]
        csv[name] = np.loadtxt(p, delimiter=",", skiprows=1, ndmin=2)
    np.savez_compressed(os.path.join(HERE, "reference_csv.npz"), **csv)

    out = {}
    # Synthetic frames: the small-stack configuration (256^2, 30 emitters) and a dense variant
    # (128^2, 60 emitters) where spots touch and several labels share components.
    small = synth_np.frames(6, 256, 256, 30, seed=11)
    dense = synth_np.frames(6, 128, 128, 60, seed=12)
    out["small"] = small
    out["dense"] = dense
    p7 = np.array([7.0])
    for name, st in (("small", small), ("dense", dense)):
        z, h, w = st.shape
        thr = R.ref_autothreshold(np.ascontiguousarray(st[0]).ctypes.data, h, w, p7.ctypes.data)
        out[f"{name}_autothreshold"] = np.array([thr])
        labels = np.zeros((2, z, h, w), np.uint32)
        regs = []
        for ri, ratio in enumerate((0.0, 10.0)):
            for f in range(z):
                img = np.ascontiguousarray(st[f])
                info = np.zeros((h * w // 4 + 2, 4))
                n = R.ref_segment(img.ctypes.data, h, w, thr, ratio, labels[ri, f].ctypes.data, info.ctypes.data,
                                  info.shape[0])
                regs.append(np.column_stack([np.full(n, ri), np.full(n, f), info[:n]]))
        out[f"{name}_labels"] = labels
        out[f"{name}_regions"] = np.concatenate(regs)
        params = np.array([7.0, 1.3, 1.3, 0.0])
        for fit in range(5):
            for ratio in (0.0, 10.0):
                cap = ((h * w) // 5 + 1) * z * 18
                buf = np.empty(cap)
                n = R.ref_localization(st.ctypes.data, buf.ctypes.data, cap, h, w, z, thr, ratio, fit, params.ctypes.data)
                out[f"{name}_loc_fit{fit}_ratio{int(ratio)}"] = buf[:n].reshape(-1, 18).copy()
        det = out[f"{name}_loc_fit0_ratio0"]
        pos = np.ascontiguousarray(np.column_stack([det[:, 0], det[:, 1], det[:, 4], det[:, 5], det[:, 15], det[:, 16]]))
        for fit in range(1, 5):
            buf = np.empty(pos.shape[0] * 18)
            n = R.ref_fitting(st.ctypes.data, pos.ctypes.data, pos.shape[0], buf.ctypes.data, buf.size, h, w, z, fit,
                              params.ctypes.data)
            out[f"{name}_fitting_fit{fit}"] = buf[:n].reshape(-1, 18).copy()
        out[f"{name}_positions"] = pos
    at = [R.ref_autothreshold(np.ascontiguousarray(stack[i]).ctypes.data, stack.shape[1], stack.shape[2], p7.ctypes.data)
          for i in range(stack.shape[0])]
    out["stack_autothreshold"] = np.array(at)
    np.savez_compressed(os.path.join(HERE, "ref_vectors.npz"), **out)
    for f in ("reference_inputs.npz", "reference_csv.npz", "ref_vectors.npz"):
        print(f, os.path.getsize(os.path.join(HERE, f)))


if __name__ == "__main__":
    main()
