"""Writes tests/golden/detect_fast.npz: outputs of the REFERENCE's own keypoint
detector (detect_keypoints_fast, core/src/preproc.cpp:436-545, through the
harness in oracle/_ref/libref_sparse.so) and of its sparse stereo tool de_pipe
(core/tools/de_pipe.cpp, DEIR pairs on stdin -> DEMR records on stdout), both
built from the reference sources by oracle/Makefile.ref.

    make -f oracle/Makefile.ref && python tests/golden/make_detect_golden.py

Inputs: the six 848x480 IR scene pairs of scenes.npz (the reference's own
article/scenes), and small synthetic images (seeded). Stored per (image,
detector config): the keypoints (x, y, response, local_std) and the reference's
DetectProfile counts; per (scene, de_pipe flags): the DEMR records (x_left,
y_left, disparity, margin). `<case>_canon` records whether the oracle's
tie-breaking order (scan position among equal responses) gives the same
keypoints as the reference, i.e. whether the case is free of deciding ties.
"""
from __future__ import annotations

import os
import struct
import subprocess
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
from oracle import detect_binding as db  # noqa: E402

DE_PIPE = os.path.join(ROOT, "oracle", "_ref", "de_pipe")
SCENES = ["night_on", "night_off", "evening_on", "evening_off", "morning_on", "morning_off"]

# name -> DetectorConfig overrides
DET_CFGS = {
    "default": {},
    "dense12": dict(cell=12, per_cell=2),
    "t12r2": dict(fast_threshold=12, nms_radius=2, cell=16, per_cell=4),
    "gw5": dict(grad_window=5, subpixel=0, min_local_std=6.0),
    "gw7r5": dict(grad_window=7, nms_radius=5, per_cell=1, border=10),
    "t4": dict(fast_threshold=4, min_response=0.0, per_cell=8, cell=24),
}

# name -> de_pipe flags
PIPE_FLAGS = {
    "default": [],
    "margin": ["--min-margin", "0.05"],
    "dense": ["--no-subpixel", "--cell", "16", "--per-cell", "2", "--max-candidates", "8"],
    "gated": ["--right-density", "6", "--min-disparity", "2", "--max-disparity", "96",
              "--fast-threshold", "10"],
}


def synthetic_images():
    rng = np.random.default_rng(2024)
    out = {}
    sq = np.full((64, 64), 30, np.uint8)
    sq[20:44, 20:44] = 220
    out["square"] = sq
    grid = np.full((128, 256), 20, np.uint8)
    for cy in range(16, 128, 32):
        for cx in range(16, 256, 32):
            grid[cy - 5:cy + 6, cx - 5:cx + 6] = 200
    out["grid"] = grid
    out["noise"] = rng.integers(0, 256, (120, 160), dtype=np.uint8)
    blobs = np.zeros((96, 200), np.float64)
    yy, xx = np.mgrid[0:96, 0:200]
    for _ in range(40):
        cx, cy, r = rng.uniform(0, 200), rng.uniform(0, 96), rng.uniform(2, 9)
        blobs += rng.uniform(-80, 80) * np.exp(-((xx - cx) ** 2 + (yy - cy) ** 2) / (2 * r * r))
    out["blobs"] = np.clip(128 + blobs + rng.normal(0, 2, blobs.shape), 0, 255).astype(np.uint8)
    return out


def deir(img, stream, frame):
    h, w = img.shape
    return b"DEIR" + struct.pack("<HHBBI", w, h, stream, 0, frame) + img.tobytes()


def run_de_pipe(pairs, flags):
    """Feed (left, right) pairs as frames 0..n-1; returns one record array per pair."""
    data = b"".join(deir(L, 1, k) + deir(R, 0, k) for k, (L, R) in enumerate(pairs))
    res = subprocess.run([DE_PIPE] + flags, input=data, capture_output=True, check=True)
    buf, pos, out = res.stdout, 0, []
    while pos < len(buf):
        assert buf[pos:pos + 4] == b"DEMR"
        w, h, frame, n = struct.unpack_from("<HHII", buf, pos + 4)
        pos += 16 + w * h
        rec = np.frombuffer(buf, np.float32, 4 * n, pos).reshape(n, 4).copy()
        pos += 16 * n
        assert frame == len(out)
        out.append(rec)
    return out


def main():
    if not os.path.exists(DE_PIPE):
        sys.exit("build the reference first: make -f oracle/Makefile.ref")
    sc = np.load(os.path.join(HERE, "scenes.npz"))
    z = {}
    images = dict(synthetic_images())
    for s in SCENES:
        images[f"{s}_L"] = sc[f"{s}_L"]
    for name, img in images.items():
        if not name.endswith("_L"):
            z[f"img_{name}"] = img
        for cname, kw in DET_CFGS.items():
            cfg = db.detector_cfg(**kw)
            k, nc, ns = db.ref_detect_keypoints_fast(img, cfg, counts=True)
            canon = db.detect_keypoints_fast(img, cfg, canonical=True)
            key = f"det_{name}_{cname}"
            z[key] = k.view(np.float32).reshape(-1, 4)
            z[key + "_counts"] = np.array([nc, ns], np.int32)
            z[key + "_canon"] = np.array(len(canon) == len(k) and bool(
                (canon.view(np.uint32) == k.view(np.uint32)).all()))
    pairs = [(sc[f"{s}_L"], sc[f"{s}_R"]) for s in SCENES]
    for fname, flags in PIPE_FLAGS.items():
        for s, rec in zip(SCENES, run_de_pipe(pairs, flags)):
            z[f"pipe_{s}_{fname}"] = rec
    path = os.path.join(HERE, "detect_fast.npz")
    np.savez_compressed(path, **z)
    print(f"wrote {path} ({os.path.getsize(path)} bytes)")
    for k in sorted(z):
        if k.endswith("_canon") and not z[k]:
            print("ties decide:", k)


if __name__ == "__main__":
    main()
