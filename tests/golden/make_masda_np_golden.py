"""Writes tests/golden/masda_np.npz: input/output vectors of the reference's
NumPy MASDA functions, imported from article/masda_stereo.py in the build
container (/root/reference). Only functions no closed form or enumeration pins
are called here: masda_sparse, masda and decide (float64 message trajectories
and the belief-ordered decision). _seg_max_excluding is pinned by enumeration
in tests/test_oracle_masda_np.py instead, and dense_masda.py cannot be
imported (it needs cv2).

    python tests/golden/make_masda_np_golden.py

Stored per problem: the inputs, the assignment as (left, right) pairs in the
dict's insertion order, and the per-iteration change history.
"""
from __future__ import annotations

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, "/root/reference/article")
import masda_stereo as ms  # noqa: E402  (reference module, read-only)


def pairs(d: dict) -> np.ndarray:
    return np.array(list(d.items()), np.int64).reshape(-1, 2)


def sparse_problems(rng):
    out = []
    for t in range(16):
        m = int(rng.integers(1, 300))
        n = int(rng.integers(1, 300))
        E = int(rng.integers(1, 4 * (m + n)))
        ei = rng.integers(0, m, E)
        ej = rng.integers(0, n, E)
        key = np.unique(ei * n + ej)            # no duplicate (i, j)
        ei, ej = key // n, key % n
        rng.shuffle(perm := np.arange(len(ei)))
        ei, ej = ei[perm], ej[perm]
        if t % 4 == 3:   # quantised scores: many exact ties
            se = (np.round(rng.uniform(-1, 1, len(ei)) * 24) / 24).astype(np.float32)
        else:
            se = rng.uniform(-1, 1, len(ei)).astype(np.float32 if t % 2 else np.float64)
        kw = dict(lam=float(rng.choice([-0.1, 0.0, 0.05])), gam=float(rng.choice([-0.1, 0.0])),
                  iters=int(rng.integers(1, 40)), damping=float(rng.choice([0.0, 0.4, 0.7])),
                  eps=float(rng.choice([1e-5, 1e-3])))
        out.append((ei.astype(np.int64), ej.astype(np.int64), se, m, n, kw))
    return out


def matrix_problems(rng):
    out = []
    for t in range(10):
        m = int(rng.integers(1, 120))
        n = int(rng.integers(1, 120))
        S = rng.uniform(-1, 1, (m, n)).astype(np.float32)
        S[rng.random((m, n)) < rng.uniform(0.5, 0.97)] = -np.inf
        kw = dict(lam=float(rng.choice([-0.1, 0.02])), gam=float(rng.choice([-0.1, 0.0])),
                  iters=int(rng.integers(1, 40)), damping=float(rng.choice([0.4, 0.5])),
                  eps=1e-5)
        out.append((S, kw))
    return out


def main():
    rng = np.random.default_rng(77)
    z = {}
    for k, (ei, ej, se, m, n, kw) in enumerate(sparse_problems(rng)):
        a, hist = ms.masda_sparse(ei, ej, se, m, n, **kw)
        z[f"sp{k}_ei"], z[f"sp{k}_ej"], z[f"sp{k}_se"] = ei, ej, se
        z[f"sp{k}_mn"] = np.array([m, n], np.int64)
        z[f"sp{k}_kw"] = np.array([kw["lam"], kw["gam"], kw["iters"], kw["damping"], kw["eps"]])
        z[f"sp{k}_assign"] = pairs(a)
        z[f"sp{k}_hist"] = hist
    for k, (S, kw) in enumerate(matrix_problems(rng)):
        a, hist = ms.masda(S, **kw)
        z[f"mx{k}_S"] = S
        z[f"mx{k}_kw"] = np.array([kw["lam"], kw["gam"], kw["iters"], kw["damping"], kw["eps"]])
        z[f"mx{k}_assign"] = pairs(a)
        z[f"mx{k}_hist"] = hist
    for k in range(6):
        m, n = int(rng.integers(1, 80)), int(rng.integers(1, 80))
        S = rng.uniform(-1, 1, (m, n)).astype(np.float32)
        S[rng.random((m, n)) < 0.6] = -np.inf
        bel = np.where(np.isfinite(S), rng.normal(0, 1, (m, n)), -np.inf)
        lam = float(rng.choice([-0.1, 0.2]))
        z[f"dc{k}_S"], z[f"dc{k}_bel"] = S, bel
        z[f"dc{k}_lam"] = np.array(lam)
        z[f"dc{k}_assign"] = pairs(ms.decide(S, bel, lam))
    path = os.path.join(HERE, "masda_np.npz")
    np.savez_compressed(path, **z)
    print(f"wrote {path} ({os.path.getsize(path)} bytes)")


if __name__ == "__main__":
    main()
