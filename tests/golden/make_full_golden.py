"""Writes tests/golden/dense_full.npz: the reference's full-volume path
(`de_dense --topk 0`, the FP32 volume and solve_row, core/tools/de_dense.cpp:
566-712) on a seeded 424x240 pair and a slanted 200x120 pair, for several flag
sets, from the reference binary built by oracle/Makefile.ref.

    make -f oracle/Makefile.ref && python tests/golden/make_full_golden.py
"""
from __future__ import annotations

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, HERE)
from make_golden import REF_BIN, run_ref  # noqa: E402

from doubleeye_amd import synth  # noqa: E402

FLAGS = {
    "it4": ["--topk", "0", "--dmax", "48", "--iters", "4", "--sigma-s", "20", "--sigma-r", "0.0392"],
    "it0": ["--topk", "0", "--dmax", "32"],
    "margin": ["--topk", "0", "--dmax", "40", "--iters", "2", "--min-margin", "0.01"],
    "nosub": ["--topk", "0", "--dmax", "48", "--iters", "3", "--no-subpixel"],
}


def main():
    if not os.path.exists(REF_BIN):
        sys.exit("build the reference first: make -f oracle/Makefile.ref")
    out = {}
    pairs = {"fronto": synth.fronto_pair(424, 240, seed=1)[:2],
             "slanted": synth.slanted_pair(200, 120, seed=4, dhi=40.0)[:2]}
    for pname, (L, R) in pairs.items():
        out[f"{pname}_L"], out[f"{pname}_R"] = L, R
        for name, flags in FLAGS.items():
            d, _, _ = run_ref(L, R, flags, conf=False)
            out[f"{pname}_{name}_disp"] = d
            out[f"{pname}_{name}_flags"] = np.array(" ".join(flags))
            print(pname, name, "coverage %.1f%%" % (100 * np.isfinite(d).mean()))
    np.savez_compressed(os.path.join(HERE, "dense_full.npz"), **out)


if __name__ == "__main__":
    main()
