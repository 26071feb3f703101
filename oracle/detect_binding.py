"""ctypes binding of the keypoint-detector oracle (oracle/liboracle.so) and of
the reference's own detector built from preproc.cpp (oracle/_ref/libref_sparse.so,
present only where /root/reference was).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py as the checker. The product package never imports
this module.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import binding
from .sparse_binding import KP_DTYPE, MATCH_DTYPE, ref_lib

# DetectorConfig field order (core/include/doubleeye/preproc.hpp:132-145)
_INT_FIELDS = ("grad_window", "cell", "per_cell", "border", "subpixel", "fast_threshold",
               "nms_radius")


class OrcDetectorCfg(C.Structure):
    _fields_ = [("grad_window", C.c_int), ("cell", C.c_int), ("per_cell", C.c_int),
                ("min_response", C.c_float), ("min_local_std", C.c_float),
                ("border", C.c_int), ("subpixel", C.c_int), ("fast_threshold", C.c_int),
                ("nms_radius", C.c_int)]


def detector_cfg(**kw) -> OrcDetectorCfg:
    c = OrcDetectorCfg()
    binding.lib().orc_detector_default_cfg(C.byref(c))
    for k, v in kw.items():
        setattr(c, k, int(v) if k in _INT_FIELDS else float(v))
    return c


def _img(img):
    a = np.ascontiguousarray(img, np.uint8)
    return a, a.ctypes.data_as(C.c_void_p)


def _cap(W, H):
    return max(16, W * H // 4)


def detect_keypoints_fast(img, cfg: OrcDetectorCfg | None = None, canonical=True,
                          counts=False):
    """Oracle detect_keypoints_fast: structured KP_DTYPE array in output order
    (and, with counts=True, the candidate and suppression-survivor counts)."""
    cfg = cfg or detector_cfg()
    a, p = _img(img)
    H, W = a.shape
    out = np.zeros(_cap(W, H), KP_DTYPE)
    n = C.c_int(0)
    cn = np.zeros(2, np.int32)
    rc = binding.lib().orc_detect_keypoints_fast(p, W, H, C.byref(cfg), int(canonical),
                                                 out.ctypes.data_as(C.c_void_p), len(out),
                                                 C.byref(n), cn.ctypes.data_as(C.c_void_p))
    if rc:
        raise ValueError(f"orc_detect_keypoints_fast rc={rc}")
    k = out[: n.value].copy()
    return (k, int(cn[0]), int(cn[1])) if counts else k


def census_at_keypoints(img, kps, hw=3, hh=3):
    """census_at (preproc.cpp:142-162) at (int(x + 0.5), int(y + 0.5)) of each
    keypoint, from the oracle census plane: 0 where the window leaves the image
    or has more than 64 bits."""
    a = np.ascontiguousarray(img, np.uint8)
    H, W = a.shape
    kps = np.ascontiguousarray(kps, KP_DTYPE)
    out = np.zeros(len(kps), np.uint64)
    if len(kps) == 0 or (2 * hw + 1) * (2 * hh + 1) - 1 > 64:
        return out
    plane = binding.census(a, hw, hh)
    x = (kps["x"] + np.float32(0.5)).astype(np.int64)
    y = (kps["y"] + np.float32(0.5)).astype(np.int64)
    inside = (x >= hw) & (y >= hh) & (x < W - hw) & (y < H - hh)
    out[inside] = plane[y[inside], x[inside]]
    return out


def is_fast_corner(img, x, y, thr) -> bool:
    a, p = _img(img)
    H, W = a.shape
    return bool(binding.lib().orc_is_fast_corner(p, W, H, int(x), int(y), int(thr)))


def fast_map(img, thr) -> np.ndarray:
    """is_fast_corner at every pixel (uint8 mask)."""
    a, p = _img(img)
    H, W = a.shape
    f = binding.lib().orc_is_fast_corner
    return np.array([[f(p, W, H, x, y, int(thr)) for x in range(W)] for y in range(H)], np.uint8)


def shi_tomasi_at(img, x, y, grad_window=3) -> float:
    a, p = _img(img)
    H, W = a.shape
    fn = binding.lib().orc_shi_tomasi_at
    fn.restype = C.c_float
    return float(fn(p, W, H, int(x), int(y), int(grad_window)))


def local_std_at(img, x, y) -> float:
    a, p = _img(img)
    H, W = a.shape
    fn = binding.lib().orc_local_std_at
    fn.restype = C.c_float
    return float(fn(p, W, H, int(x), int(y)))


def refine_disparity(L, R, kl, kr, matches, half=3):
    """Oracle refine_disparity on a copy of the matches; returns the copy."""
    a, pa = _img(L)
    b, pb = _img(R)
    H, W = a.shape
    kl = np.ascontiguousarray(kl, KP_DTYPE)
    kr = np.ascontiguousarray(kr, KP_DTYPE)
    m = np.ascontiguousarray(matches, MATCH_DTYPE).copy()
    binding.lib().orc_refine_disparity(pa, pb, W, H, kl.ctypes.data_as(C.c_void_p), len(kl),
                                       kr.ctypes.data_as(C.c_void_p), len(kr),
                                       m.ctypes.data_as(C.c_void_p), len(m), int(half))
    return m


# --- the reference itself -------------------------------------------------------

def _cfg_arrays(cfg: OrcDetectorCfg):
    ic = np.array([cfg.grad_window, cfg.cell, cfg.per_cell, cfg.border, cfg.subpixel,
                   cfg.fast_threshold, cfg.nms_radius], np.int32)
    fc = np.array([cfg.min_response, cfg.min_local_std], np.float32)
    return ic, fc


def _ref_detect(fn, img, cfg, extra=()):
    cfg = cfg or detector_cfg()
    a, p = _img(img)
    H, W = a.shape
    ic, fc = _cfg_arrays(cfg)
    out = np.zeros(_cap(W, H), KP_DTYPE)
    n = C.c_int(0)
    rc = fn(p, W, H, ic.ctypes.data_as(C.c_void_p), fc.ctypes.data_as(C.c_void_p),
            out.ctypes.data_as(C.c_void_p), len(out), C.byref(n), *extra)
    if rc:
        raise ValueError(f"reference detector rc={rc}")
    return out[: n.value].copy()


def ref_detect_keypoints_fast(img, cfg: OrcDetectorCfg | None = None, counts=False):
    """The reference's detect_keypoints_fast (and, with counts=True, its
    DetectProfile candidates and suppressed)."""
    cn = np.zeros(2, np.int32)
    k = _ref_detect(ref_lib().ref_detect_keypoints_fast, img, cfg,
                    (cn.ctypes.data_as(C.c_void_p),))
    return (k, int(cn[0]), int(cn[1])) if counts else k


def ref_detect_keypoints_dense(img, cfg: OrcDetectorCfg | None = None):
    return _ref_detect(ref_lib().ref_detect_keypoints_dense, img, cfg)


def ref_is_fast_corner(img, x, y, thr) -> bool:
    a, p = _img(img)
    H, W = a.shape
    return bool(ref_lib().ref_is_fast_corner(p, W, H, int(x), int(y), int(thr)))


def ref_shi_tomasi_maps(img, grad_window=3):
    """(shi_tomasi_at at every pixel, dense shi_tomasi_response)."""
    a, p = _img(img)
    H, W = a.shape
    at = np.zeros((H, W), np.float32)
    dense = np.zeros((H, W), np.float32)
    ref_lib().ref_shi_tomasi_maps(p, W, H, int(grad_window), at.ctypes.data_as(C.c_void_p),
                                  dense.ctypes.data_as(C.c_void_p))
    return at, dense


def ref_refine_disparity(L, R, kl, kr, matches, half=3):
    a, pa = _img(L)
    b, pb = _img(R)
    H, W = a.shape
    kl = np.ascontiguousarray(kl, KP_DTYPE)
    kr = np.ascontiguousarray(kr, KP_DTYPE)
    m = np.ascontiguousarray(matches, MATCH_DTYPE).copy()
    ref_lib().ref_refine_disparity(pa, pb, W, H, kl.ctypes.data_as(C.c_void_p), len(kl),
                                   kr.ctypes.data_as(C.c_void_p), len(kr),
                                   m.ctypes.data_as(C.c_void_p), len(m), int(half))
    return m


def cell_occupancy(kps, width, height, cell) -> float:
    """cell_occupancy (preproc.cpp:549-564)."""
    if width <= 0 or height <= 0:
        return 0.0
    cell = max(4, cell)
    cols, rows = (width + cell - 1) // cell, (height + cell - 1) // cell
    hit = np.zeros((rows, cols), bool)
    for k in kps:
        cx, cy = int(k["x"]) // cell, int(k["y"]) // cell
        if 0 <= cx < cols and 0 <= cy < rows:
            hit[cy, cx] = True
    return float(hit.sum()) / float(hit.size)


def pipe_match_cfg(**kw):
    """de_pipe's matcher settings (core/tools/de_pipe.cpp:42-45)."""
    from .sparse_binding import match_cfg
    base = dict(lam=-0.1, gamma=-0.1, iterations=20, damping=0.4)
    base.update(kw)
    return match_cfg(**base)


def pipeline(L, R, det=None, det_right=None, census=(3, 3), mcfg=None, min_margin=0.0,
             subpixel=True, canonical=True):
    """core/tools/de_pipe.cpp:167-199 on the oracle: (kl, kr, matches, records)."""
    from .sparse_binding import match_masda
    det = det or detector_cfg()
    det_right = det_right or det
    mcfg = mcfg or pipe_match_cfg()
    kl = detect_keypoints_fast(L, det, canonical)
    kr = detect_keypoints_fast(R, det_right, canonical)
    dl = census_at_keypoints(L, kl, *census)
    dr = census_at_keypoints(R, kr, *census)
    m, _ = match_masda(kl, dl, kr, dr, mcfg, canonical=canonical)
    if min_margin > 0:
        m = m[~(m["margin"] < np.float32(min_margin))]
    if subpixel:
        m = refine_disparity(L, R, kl, kr, m)
    rec = np.zeros((len(m), 4), np.float32)
    if len(m):
        rec[:, 0] = kl["x"][m["left"]]
        rec[:, 1] = kl["y"][m["left"]]
        rec[:, 2] = m["disparity"]
        rec[:, 3] = m["margin"]
    return kl, kr, m, rec


def pipe_args(flags):
    """de_pipe command-line flags (de_pipe.cpp:50-62) -> pipeline() keywords."""
    det, mc = {}, {}
    kw = dict(min_margin=0.0, subpixel=True)
    right_density = 0
    i = 0
    while i < len(flags):
        a = flags[i]
        if a == "--no-subpixel":
            kw["subpixel"] = False
            i += 1
            continue
        v = flags[i + 1]
        if a == "--right-density":
            right_density = int(v)
        elif a == "--min-margin":
            kw["min_margin"] = float(v)
        elif a == "--fast-threshold":
            det["fast_threshold"] = int(v)
        elif a == "--max-disparity":
            mc["max_disparity"] = float(v)
        elif a == "--min-disparity":
            mc["min_disparity"] = float(v)
        elif a == "--cell":
            det["cell"] = int(v)
        elif a == "--per-cell":
            det["per_cell"] = int(v)
        elif a == "--max-candidates":
            mc["max_candidates"] = int(v)
        else:
            raise ValueError(a)
        i += 2
    dr = dict(det)
    if right_density > 0:
        dr["per_cell"] = right_density
    return det, dr, mc, kw
