"""Writes tests/golden/dense_prior.npz: the reference's `de_dense --prior` and
`de_dense --c2f` (core/tools/de_dense.cpp:797-853, 1156-1296, 1771-1966) on
seeded synthetic pairs and one IR scene pair, from the reference binary built
by oracle/Makefile.ref (`--threads 1`, its deterministic configuration).

    make -f oracle/Makefile.ref && python tests/golden/make_prior_golden.py

Stored per case: the inputs (synthetic pairs by seed, the scene by name), the
prior given to `--prior` (full or half resolution), the flags, the disparity
plane in full and the six `--out-conf` planes as a sha256 digest.
"""
from __future__ import annotations

import hashlib
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, HERE)
from make_golden import REF_BIN, run_ref  # noqa: E402

from doubleeye_amd import synth  # noqa: E402


def priors(gt, rng):
    """A noisy full-resolution prior, a half-resolution one (half the
    disparity, nearest samples) and a constant one."""
    H, W = gt.shape
    g = np.where(np.isfinite(gt), gt, np.nanmedian(gt)).astype(np.float32)
    full = (g + rng.normal(0, 1.2, g.shape)).astype(np.float32)
    half = (g[::2, ::2][: H // 2, : W // 2] / 2 + rng.normal(0, 0.8, (H // 2, W // 2)))
    const = np.full_like(g, float(np.nanmedian(g)))
    return {"full": full, "half": half.astype(np.float32), "const": const}


CASES = [
    # (case, pair, prior kind or None, flags)
    ("fr_full_b2", "fronto", "full", ["--dmax", "48"]),
    ("fr_half_b4_lrc", "fronto", "half", ["--dmax", "48", "--band", "4", "--lrc", "--iters", "3"]),
    ("fr_const_b6", "fronto", "const", ["--dmax", "40", "--band", "6", "--min-margin", "0.01"]),
    ("fr_c2f", "fronto", None, ["--dmax", "48", "--c2f"]),
    ("sl_c2f_b3_lrc", "slanted", None, ["--dmax", "64", "--c2f", "--band", "3", "--lrc",
                                         "--iters", "4", "--sigma-s", "20",
                                         "--sigma-r", "0.0392"]),
    ("sl_full_b1", "slanted", "full", ["--dmax", "64", "--band", "1", "--fit-parabola"]),
    ("ir_c2f", "scene", None, ["--dmax", "64", "--c2f"]),
]


def main():
    if not os.path.exists(REF_BIN):
        sys.exit("build the reference first: make -f oracle/Makefile.ref")
    rng = np.random.default_rng(11)
    fl, fr, fgt = synth.fronto_pair(424, 240, seed=1)
    sl, sr, sgt = synth.slanted_pair(200, 120, seed=4, dhi=40.0)[:3]
    sc = np.load(os.path.join(HERE, "scenes.npz"))
    pairs = {"fronto": (fl, fr, fgt), "slanted": (sl, sr, sgt),
             "scene": (sc["night_on_L"], sc["night_on_R"], None)}
    pri = {k: priors(v[2], rng) for k, v in pairs.items() if v[2] is not None}
    out = {}
    for case, pair, kind, flags in CASES:
        L, R, _ = pairs[pair]
        f = list(flags)
        if kind is not None:
            p = pri[pair][kind]
            path = os.path.join("/tmp", f"prior_{case}.f32")
            p.tofile(path)
            f += ["--prior", path]
            out[f"{case}_prior"] = p
        d, c, _ = run_ref(L, R, f)
        out[f"{case}_pair"] = np.array(pair)
        out[f"{case}_flags"] = np.array(" ".join(flags))
        out[f"{case}_disp"] = d
        out[f"{case}_cand_sha256"] = np.array(hashlib.sha256(c.tobytes()).hexdigest())
        print(case, "coverage %.1f%%" % (100 * np.isfinite(d).mean()))
    path = os.path.join(HERE, "dense_prior.npz")
    np.savez_compressed(path, **out)
    print(f"wrote {path} ({os.path.getsize(path)} bytes)")


if __name__ == "__main__":
    main()
