"""Shared cases of the keypoint-detector tests: the golden fixture written from
the reference (tests/golden/make_detect_golden.py), its images and configs."""
from __future__ import annotations

import os
import sys

import numpy as np

from conftest import GOLDEN

sys.path.insert(0, GOLDEN)
from make_detect_golden import DET_CFGS, PIPE_FLAGS, SCENES  # noqa: E402,F401

_Z = None
_S = None


def golden():
    global _Z
    if _Z is None:
        _Z = np.load(os.path.join(GOLDEN, "detect_fast.npz"))
    return _Z


def scenes():
    global _S
    if _S is None:
        _S = np.load(os.path.join(GOLDEN, "scenes.npz"))
    return _S


def image(name: str) -> np.ndarray:
    if name.endswith("_L"):
        return scenes()[name]
    return golden()[f"img_{name}"]


def image_names():
    z = golden()
    return [k[4:] for k in z.files if k.startswith("img_")] + [f"{s}_L" for s in SCENES]


def det_cases():
    """(image name, config name) for every golden detector case."""
    return [(i, c) for i in image_names() for c in DET_CFGS]


def det_golden(img_name: str, cfg_name: str):
    z = golden()
    key = f"det_{img_name}_{cfg_name}"
    return z[key], tuple(int(v) for v in z[key + "_counts"]), bool(z[key + "_canon"])


def as_rows(kps) -> np.ndarray:
    """Keypoints as an (n, 4) float32 array."""
    a = np.ascontiguousarray(kps)
    if a.dtype.names:
        return a.view(np.float32).reshape(-1, 4)
    return np.asarray(a, np.float32).reshape(-1, 4)


def same_bits(a, b) -> bool:
    a, b = np.ascontiguousarray(a, np.float32), np.ascontiguousarray(b, np.float32)
    return a.shape == b.shape and bool((a.view(np.uint32) == b.view(np.uint32)).all())
