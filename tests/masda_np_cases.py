"""Golden problems of tests/golden/masda_np.npz (written from the reference's
article/masda_stereo.py) and helpers shared by the oracle and GPU tests."""
from __future__ import annotations

import os

import numpy as np

from conftest import GOLDEN


def load():
    z = np.load(os.path.join(GOLDEN, "masda_np.npz"), allow_pickle=False)
    sp, mx, dc = [], [], []
    k = 0
    while f"sp{k}_ei" in z.files:
        lam, gam, iters, damping, eps = z[f"sp{k}_kw"]
        m, n = (int(v) for v in z[f"sp{k}_mn"])
        sp.append(dict(ei=z[f"sp{k}_ei"], ej=z[f"sp{k}_ej"], se=z[f"sp{k}_se"], m=m, n=n,
                       kw=dict(lam=float(lam), gam=float(gam), iters=int(iters),
                               damping=float(damping), eps=float(eps)),
                       assign=z[f"sp{k}_assign"], hist=z[f"sp{k}_hist"]))
        k += 1
    k = 0
    while f"mx{k}_S" in z.files:
        lam, gam, iters, damping, eps = z[f"mx{k}_kw"]
        mx.append(dict(S=z[f"mx{k}_S"], kw=dict(lam=float(lam), gam=float(gam), iters=int(iters),
                                                damping=float(damping), eps=float(eps)),
                       assign=z[f"mx{k}_assign"], hist=z[f"mx{k}_hist"]))
        k += 1
    k = 0
    while f"dc{k}_S" in z.files:
        dc.append(dict(S=z[f"dc{k}_S"], bel=z[f"dc{k}_bel"], lam=float(z[f"dc{k}_lam"]),
                       assign=z[f"dc{k}_assign"]))
        k += 1
    return sp, mx, dc


def pairs(d: dict) -> np.ndarray:
    return np.array(list(d.items()), np.int64).reshape(-1, 2)


def as_set(p) -> set:
    return {(int(a), int(b)) for a, b in np.asarray(p).reshape(-1, 2)}


def has_belief_ties(belief) -> bool:
    b = np.asarray(belief)
    b = b[np.isfinite(b)]
    return len(np.unique(b)) < len(b)
