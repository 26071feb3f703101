"""Writes tests/golden/sparse_match.npz from the REFERENCE's own sparse
matcher: core/src/match.cpp compiled from the reference sources by
oracle/Makefile.ref into oracle/_ref/libref_sparse.so (a C harness over
doubleeye::match_masda, match_mutual_nn and match_brute_force).

Run in the build container (where /root/reference exists):

    make -f oracle/Makefile.ref && python tests/golden/make_sparse_golden.py

For every problem of tests/sparse_cases.py the fixture holds the inputs
(keypoint x/y, descriptors, configuration, ratio) and the reference's outputs:
match_masda matches and stats, match_mutual_nn matches, and for the tiny
brute-force problems the exhaustive optimum.
"""
from __future__ import annotations

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(HERE))
import sparse_cases as sc  # noqa: E402
from oracle import sparse_binding as sb  # noqa: E402

STATS_FIELDS = ["left", "right", "candidates", "iterations_run", "converged",
                "final_max_delta", "oscillating"]


def orc_cfg(c: dict):
    kw = {("lam" if k == "lam" else k): v for k, v in c.items()}
    return sb.match_cfg(**kw)


def main():
    cases = sc.small_cases() + sc.brute_force_cases() + sc.random_cases() + sc.large_cases()
    out = {"names": np.array([p["name"] for p in cases])}
    for k, p in enumerate(cases):
        c = orc_cfg(p["cfg"])
        m, st = sb.ref_match_masda(p["L"], p["dl"], p["R"], p["dr"], c)
        nn, _ = sb.ref_match_mutual_nn(p["L"], p["dl"], p["R"], p["dr"], c, p["ratio"])
        out[f"{k}_L"] = p["L"]
        out[f"{k}_dl"] = p["dl"]
        out[f"{k}_R"] = p["R"]
        out[f"{k}_dr"] = p["dr"]
        out[f"{k}_cfg"] = sc.cfg_array(p["cfg"])
        out[f"{k}_ratio"] = np.float32(p["ratio"])
        out[f"{k}_masda"] = m
        out[f"{k}_stats"] = np.array([float(st[f]) for f in STATS_FIELDS], np.float64)
        out[f"{k}_nn"] = nn
        if p["name"].startswith("brute_") or p["name"] == "degenerate":
            out[f"{k}_brute"] = sb.ref_match_brute_force(p["L"], p["dl"], p["R"], p["dr"], c)
    path = os.path.join(HERE, "sparse_match.npz")
    np.savez_compressed(path, **out)
    print(f"wrote {path}: {len(cases)} problems, {os.path.getsize(path)} bytes")


if __name__ == "__main__":
    main()
