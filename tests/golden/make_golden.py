"""Writes the golden fixtures under tests/golden/ from the REFERENCE's own CPU
tool, `de_dense --threads 1` (its only deterministic configuration), built from
the reference sources by oracle/Makefile.ref into oracle/_ref/.

Run in the build container (where /root/reference exists):

    make -f oracle/Makefile.ref && python tests/golden/make_golden.py

Fixtures written (data only: inputs and the reference's outputs):
  scenes.npz        the six 848x480 IR pairs held by the reference
                    (article/scenes/*/{left,right}.y8) with their meta.txt
                    coverage_pct known answers, plus sha256 of de_dense
                    --dmax 64 --out for each
  dense_c1.npz      seeded 424x240 fronto pair; de_dense outputs (disp and the
                    six --out-conf planes) in full for the C1 flags and as
                    sha256 digests for the other flag sets
  volume_f32.npz    small pair and de_dense --dump-vol (FP32 full volume)
"""
from __future__ import annotations

import hashlib
import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
from doubleeye_amd import synth  # noqa: E402

REF_BIN = os.path.join(ROOT, "oracle", "_ref", "de_dense")
SCENES = "/root/reference/article/scenes"
SCENE_NAMES = ["night_on", "night_off", "evening_on", "evening_off",
               "morning_on", "morning_off"]

# (name, de_dense flags) for the C1 fixture
C1_FLAGS = {
    "c1": ["--dmax", "48", "--iters", "5", "--sigma-s", "20", "--sigma-r", "0.0392"],
    "default48": ["--dmax", "48"],
    "lrc_it2": ["--dmax", "48", "--iters", "2", "--lrc"],
    "nosub_margin": ["--dmax", "40", "--no-subpixel", "--min-margin", "0.01"],
    "parabola_ratio": ["--dmax", "48", "--fit-parabola", "--iters", "3",
                       "--min-ratio", "1.05"],
    "d96_it8": ["--dmax", "96", "--iters", "8", "--sigma-s", "20",
                "--sigma-r", "0.0392", "--lrc"],
}


def run_ref(L, R, flags, conf=True, vol=False):
    H, W = L.shape
    with tempfile.TemporaryDirectory() as td:
        lp, rp = os.path.join(td, "l.y8"), os.path.join(td, "r.y8")
        L.tofile(lp)
        R.tofile(rp)
        op, cp, vp = (os.path.join(td, n) for n in ("o.f32", "c.f32", "v.f32"))
        cmd = [REF_BIN, lp, rp, str(W), str(H), "--threads", "1", "--out", op] + flags
        if conf:
            cmd += ["--out-conf", cp]
        if vol:
            cmd += ["--dump-vol", vp]
        subprocess.run(cmd, check=True, capture_output=True)
        d = np.fromfile(op, np.float32).reshape(H, W)
        c = np.fromfile(cp, np.float32).reshape(6, H, W) if conf else None
        v = np.fromfile(vp, np.float32) if vol else None
    return d, c, v


def main():
    if not os.path.exists(REF_BIN):
        sys.exit("build the reference first: make -f oracle/Makefile.ref")
    # --- scenes: inputs, known answers, hashes of the reference output ---
    out = {}
    for sc in SCENE_NAMES:
        L = np.fromfile(f"{SCENES}/{sc}/left.y8", np.uint8).reshape(480, 848)
        R = np.fromfile(f"{SCENES}/{sc}/right.y8", np.uint8).reshape(480, 848)
        meta = dict(line.split(None, 1) for line in
                    open(f"{SCENES}/{sc}/meta.txt").read().splitlines() if line.strip())
        d, _, _ = run_ref(L, R, ["--dmax", "64"], conf=False)
        out[f"{sc}_L"] = L
        out[f"{sc}_R"] = R
        out[f"{sc}_coverage_pct"] = np.float64(meta["coverage_pct"])
        out[f"{sc}_disp_sha256"] = np.array(hashlib.sha256(d.tobytes()).hexdigest())
        print(sc, "coverage", meta["coverage_pct"],
              "ref %.1f" % (100 * np.isfinite(d).mean()))
    np.savez_compressed(os.path.join(HERE, "scenes.npz"), **out)

    # --- C1 seeded synthetic pair ---
    L, R, _ = synth.fronto_pair(424, 240, seed=1)
    out = {"L": L, "R": R}
    for name, flags in C1_FLAGS.items():
        d, c, _ = run_ref(L, R, flags)
        if name == "c1":   # full arrays for the headline C1 flags only
            out[f"{name}_disp"] = d
            out[f"{name}_cand"] = c
        out[f"{name}_disp_sha256"] = np.array(hashlib.sha256(d.tobytes()).hexdigest())
        out[f"{name}_cand_sha256"] = np.array(hashlib.sha256(c.tobytes()).hexdigest())
        out[f"{name}_flags"] = np.array(" ".join(flags))
    np.savez_compressed(os.path.join(HERE, "dense_c1.npz"), **out)

    # --- FP32 full volume (--dump-vol), small pair ---
    L, R, _ = synth.fronto_pair(120, 72, seed=3, disparities=(3, 6, 10))
    flags = ["--dmax", "16", "--sigma-s", "20", "--sigma-r", "0.0392"]
    d, _, v = run_ref(L, R, flags, conf=False, vol=True)
    np.savez_compressed(os.path.join(HERE, "volume_f32.npz"), L=L, R=R,
                        vol=v.reshape(72, 120, 16), disp=d,
                        flags=np.array(" ".join(flags)))
    print("wrote fixtures to", HERE)


if __name__ == "__main__":
    main()
