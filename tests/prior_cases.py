"""Cases of the --prior / --c2f fixture (tests/golden/make_prior_golden.py)."""
from __future__ import annotations

import hashlib
import os

import numpy as np

from conftest import GOLDEN, flags_to_kwargs

_Z = None


def golden():
    global _Z
    if _Z is None:
        _Z = np.load(os.path.join(GOLDEN, "dense_prior.npz"))
    return _Z


def pair(name):
    from doubleeye_amd import synth
    if name == "fronto":
        return synth.fronto_pair(424, 240, seed=1)[:2]
    if name == "slanted":
        return synth.slanted_pair(200, 120, seed=4, dhi=40.0)[:2]
    z = np.load(os.path.join(GOLDEN, "scenes.npz"))
    return z["night_on_L"], z["night_on_R"]


def cases():
    z = golden()
    return sorted(k[: -len("_flags")] for k in z.files if k.endswith("_flags"))


def case(name):
    """(L, R, cfg kwargs, band, c2f, prior or None, ref disp, ref cand sha256)."""
    z = golden()
    kw = flags_to_kwargs(str(z[f"{name}_flags"]))
    band = kw.pop("band", 2)
    c2f = bool(kw.pop("c2f", 0))
    L, R = pair(str(z[f"{name}_pair"]))
    prior = z[f"{name}_prior"] if f"{name}_prior" in z.files else None
    return L, R, kw, band, c2f, prior, z[f"{name}_disp"], str(z[f"{name}_cand_sha256"])


def sha(a) -> str:
    return hashlib.sha256(np.ascontiguousarray(a, np.float32).tobytes()).hexdigest()
