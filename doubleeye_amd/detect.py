"""FAST keypoint detection and the detect -> match pipeline on the MI355X.

Mirrors core/include/doubleeye/preproc.hpp and core/tools/de_pipe.cpp:

* ``DetectorConfig`` -- ``doubleeye::DetectorConfig`` (preproc.hpp:132-145),
  same field names and defaults.
* ``detect_keypoints_fast(img, cfg, prof)`` -- ``doubleeye::detect_keypoints_fast``
  (preproc.hpp:202-204); ``prof`` is a ``DetectProfile`` (preproc.hpp:191-198).
* ``is_fast_corner`` / ``shi_tomasi_at`` (preproc.hpp:179, 183) through
  ``fast_score_map``, which evaluates both at every pixel.
* ``census_at_keypoints`` -- ``census_at`` (preproc.hpp:96-97) at each keypoint's
  rounded position, as de_pipe.cpp:168-172 computes the descriptors.
* ``refine_disparity`` (match.hpp:293-296).
* ``cell_occupancy`` (preproc.hpp:209-210), a host-side summary.
* ``sparse_pipeline(left, right, ...)`` -- one stereo pair through de_pipe's
  detect -> describe -> match_masda -> margin filter -> refine chain
  (de_pipe.cpp:167-199), all on the device.

Keypoints are ``KEYPOINT`` structured arrays, matches ``MATCH`` arrays (both
from ``doubleeye_amd.sparse``).
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, fields

import numpy as np

from ._lib import DeDetectFrameC, DeDetectorCfgC, DeDetectProfileC, DeMatchStatsC, check, lib
from .sparse import KEYPOINT, MATCH, MatchConfig, MatchStats, keypoints

__all__ = ["DetectorConfig", "DetectProfile", "detect_keypoints_fast", "fast_score_map",
           "is_fast_corner", "shi_tomasi_at", "census_at_keypoints", "refine_disparity",
           "cell_occupancy", "max_keypoints", "sparse_pipeline", "pipe_match_config",
           "DetectFrame", "detect_keypoints_fast_batch_dev"]


@dataclass
class DetectorConfig:
    grad_window: int = 3        # structure-tensor box window
    cell: int = 32              # grid cell for spatial distribution
    per_cell: int = 3           # keep top-N per cell
    min_response: float = 1.0
    min_local_std: float = 2.0  # texture floor
    border: int = 8
    subpixel: bool = True
    fast_threshold: int = 8
    nms_radius: int = 3         # candidate suppression radius

    def to_c(self) -> DeDetectorCfgC:
        c = DeDetectorCfgC()
        for f in fields(self):
            v = getattr(self, f.name)
            setattr(c, f.name, int(v) if f.name == "subpixel" else v)
        return c


@dataclass
class DetectProfile:
    """``doubleeye::DetectProfile``: stage times (GPU events) and counts."""
    fast_ms: float = 0.0
    nms_ms: float = 0.0
    refine_ms: float = 0.0
    candidates: int = 0
    suppressed: int = 0

    def total_ms(self) -> float:
        return self.fast_ms + self.nms_ms + self.refine_ms


def _ctx(ctx):
    if ctx is None:
        from .dense import default_context
        ctx = default_context()
    return ctx


def _img(img) -> np.ndarray:
    a = np.ascontiguousarray(img, np.uint8)
    if a.ndim != 2:
        raise ValueError("expected a 2-D uint8 image")
    return a


def max_keypoints(W: int, H: int, cfg: DetectorConfig | None = None) -> int:
    """Upper bound on the keypoints one W x H image yields."""
    return int(lib().de_detect_max_keypoints(W, H, C.byref((cfg or DetectorConfig()).to_c())))


def detect_keypoints_fast(img, cfg: DetectorConfig | None = None,
                          prof: DetectProfile | None = None, ctx=None) -> np.ndarray:
    """Keypoints in row-major cell order, each cell's by descending response."""
    a = _img(img)
    H, W = a.shape
    cfg = cfg or DetectorConfig()
    cap = max(1, max_keypoints(W, H, cfg))
    out = np.zeros(cap, KEYPOINT)
    n = C.c_int(0)
    p = DeDetectProfileC()
    check(lib().de_detect_keypoints_fast(_ctx(ctx).h, a.ctypes.data_as(C.c_void_p), W, H,
                                         C.byref(cfg.to_c()), out.ctypes.data_as(C.c_void_p),
                                         cap, C.byref(n), C.byref(p) if prof else None))
    if prof is not None:
        for name, _ in DeDetectProfileC._fields_:
            setattr(prof, name, getattr(p, name))
    return out[: n.value].copy()


@dataclass
class DetectFrame:
    """Device buffers of one image of a detector batch (DeviceArray objects):
    ``img`` (H, W) uint8, ``out`` KEYPOINT with at least max_keypoints(W, H, cfg)
    entries, ``n_out`` one int32."""
    img: object
    out: object
    n_out: object

    def to_c(self) -> DeDetectFrameC:
        H, W = self.img.shape
        return DeDetectFrameC(self.img.ptr, W, H, self.out.ptr, self.out.shape[0],
                              self.n_out.ptr)


def detect_keypoints_fast_batch_dev(ctx, frames, cfgs, stream: int = 0):
    """``detect_keypoints_fast`` over device-resident images in one chain of
    launches; results land in each frame's ``out`` / ``n_out``."""
    n = len(frames)
    fa = (DeDetectFrameC * n)(*[f.to_c() for f in frames])
    ca = (DeDetectorCfgC * n)(*[c.to_c() for c in cfgs])
    check(lib().de_detect_keypoints_fast_batch_dev(ctx.h, fa, ca, n, C.c_void_p(stream or None)))


def fast_score_map(img, threshold: int = 8, grad_window: int = 3, ctx=None):
    """(is_fast_corner mask as uint8, shi_tomasi_at response) at every pixel."""
    a = _img(img)
    H, W = a.shape
    corner = np.zeros((H, W), np.uint8)
    resp = np.zeros((H, W), np.float32)
    check(lib().de_fast_score_map(_ctx(ctx).h, a.ctypes.data_as(C.c_void_p), W, H,
                                  int(threshold), int(grad_window),
                                  corner.ctypes.data_as(C.c_void_p),
                                  resp.ctypes.data_as(C.c_void_p)))
    return corner, resp


def is_fast_corner(img, x: int, y: int, threshold: int, ctx=None) -> bool:
    corner, _ = fast_score_map(img, threshold, 3, ctx)
    return bool(corner[y, x])


def shi_tomasi_at(img, x: int, y: int, grad_window: int = 3, ctx=None) -> float:
    _, resp = fast_score_map(img, 8, grad_window, ctx)
    return float(resp[y, x])


def census_at_keypoints(img, kps, hw: int = 3, hh: int = 3, ctx=None) -> np.ndarray:
    """census_at at (int(x + 0.5), int(y + 0.5)) of each keypoint (0 outside)."""
    a = _img(img)
    H, W = a.shape
    k = keypoints(kps)
    out = np.zeros(len(k), np.uint64)
    check(lib().de_census_at_keypoints(_ctx(ctx).h, a.ctypes.data_as(C.c_void_p), W, H, hw, hh,
                                       k.ctypes.data_as(C.c_void_p) if len(k) else None, len(k),
                                       out.ctypes.data_as(C.c_void_p)))
    return out


def refine_disparity(left, right, kl, kr, matches, half: int = 3, ctx=None) -> np.ndarray:
    """Sub-pixel disparity from the SAD parabola; returns refined copies."""
    a, b = _img(left), _img(right)
    H, W = a.shape
    kl, kr = keypoints(kl), keypoints(kr)
    m = np.ascontiguousarray(matches, MATCH).copy()
    check(lib().de_refine_disparity(_ctx(ctx).h, a.ctypes.data_as(C.c_void_p),
                                    b.ctypes.data_as(C.c_void_p), W, H,
                                    kl.ctypes.data_as(C.c_void_p) if len(kl) else None, len(kl),
                                    kr.ctypes.data_as(C.c_void_p) if len(kr) else None, len(kr),
                                    m.ctypes.data_as(C.c_void_p) if len(m) else None, len(m),
                                    int(half)))
    return m


def cell_occupancy(kps, width: int, height: int, cell: int) -> float:
    """Fraction of cells holding at least one keypoint (preproc.cpp:549-564)."""
    if width <= 0 or height <= 0:
        return 0.0
    cell = max(4, cell)
    cols, rows = (width + cell - 1) // cell, (height + cell - 1) // cell
    k = keypoints(kps)
    cx = k["x"].astype(np.int64) // cell
    cy = k["y"].astype(np.int64) // cell
    ok = (cx >= 0) & (cx < cols) & (cy >= 0) & (cy < rows)
    hit = np.zeros(rows * cols, bool)
    hit[cy[ok] * cols + cx[ok]] = True
    return float(hit.sum()) / float(hit.size)


def pipe_match_config(**kw) -> MatchConfig:
    """de_pipe's matcher settings (de_pipe.cpp:42-45): lambda = gamma = -0.1,
    20 iterations, damping 0.4, over MatchConfig's other defaults."""
    c = MatchConfig(lam=-0.1, gamma=-0.1, iterations=20, damping=0.4)
    for k, v in kw.items():
        setattr(c, k, v)
    return c


def sparse_pipeline(left, right, det: DetectorConfig | None = None,
                    det_right: DetectorConfig | None = None, census=(3, 3),
                    cfg: MatchConfig | None = None, min_margin: float = 0.0,
                    subpixel: bool = True, stats: MatchStats | None = None, ctx=None):
    """One pair through de_pipe: returns (left keypoints, right keypoints,
    matches, records) with records the DEMR (x, y, disparity, margin) rows."""
    a, b = _img(left), _img(right)
    if a.shape != b.shape:
        raise ValueError("left and right must have one shape")
    H, W = a.shape
    det = det or DetectorConfig()
    dr = det_right or det
    cfg = cfg or pipe_match_config()
    cl, cr = max(1, max_keypoints(W, H, det)), max(1, max_keypoints(W, H, dr))
    kl = np.zeros(cl, KEYPOINT)
    kr = np.zeros(cr, KEYPOINT)
    cap = max(1, min(cl, cr))
    out = np.zeros(cap, MATCH)
    rec = np.zeros((cap, 4), np.float32)
    nl, nr, n = C.c_int(0), C.c_int(0), C.c_int(0)
    st = DeMatchStatsC()
    check(lib().de_sparse_pipeline(
        _ctx(ctx).h, a.ctypes.data_as(C.c_void_p), b.ctypes.data_as(C.c_void_p), W, H,
        C.byref(det.to_c()), C.byref(dr.to_c()), int(census[0]), int(census[1]),
        C.byref(cfg.to_c()), C.c_float(min_margin), int(bool(subpixel)),
        kl.ctypes.data_as(C.c_void_p), cl, C.byref(nl), kr.ctypes.data_as(C.c_void_p), cr,
        C.byref(nr), out.ctypes.data_as(C.c_void_p), rec.ctypes.data_as(C.c_void_p), cap,
        C.byref(n), C.byref(st)))
    if stats is not None:
        stats._fill(st)
    return (kl[: nl.value].copy(), kr[: nr.value].copy(), out[: n.value].copy(),
            rec[: n.value].copy())
