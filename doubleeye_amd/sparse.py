"""Sparse keypoint MASDA on the MI355X through the C ABI.

Mirrors core/include/doubleeye/match.hpp:

* ``MatchConfig`` -- ``doubleeye::MatchConfig`` (match.hpp:136-176), same field
  names and defaults (``lam`` for ``lambda``); the prior and smoothness prior
  are passed to the calls, as dicts of arrays.
* ``match_masda(left, desc_left, right, desc_right, cfg, stats)`` --
  ``doubleeye::match_masda`` (match.hpp:221-226).
* ``match_mutual_nn(..., ratio)`` -- ``doubleeye::match_mutual_nn``
  (match.hpp:228-235).
* ``matching_objective(matches, cfg, n_left, n_right)`` (match.hpp:300-303).

Keypoints are numpy structured arrays of ``KEYPOINT`` (``doubleeye::Keypoint``)
or (N, 2) x/y arrays; descriptors are uint64. Matches come back as a
structured array of ``MATCH`` (``doubleeye::Match``) in the reference's order.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, fields

import numpy as np

from ._lib import (DeMatchCfgC, DeMatchStatsC, DePriorC, DeSmoothC, DeSparseFrameC, check,
                   lib)

__all__ = ["KEYPOINT", "MATCH", "MatchConfig", "MatchStats", "keypoints", "match_masda",
           "match_mutual_nn", "match_candidates", "matching_objective", "SparseFrame",
           "match_masda_batch_dev"]

KEYPOINT = np.dtype([("x", "<f4"), ("y", "<f4"), ("response", "<f4"), ("local_std", "<f4")])
MATCH = np.dtype([("left", "<i4"), ("right", "<i4"), ("score", "<f4"), ("belief", "<f4"),
                  ("disparity", "<f4"), ("dy", "<f4"), ("margin", "<f4")])
STATS = np.dtype([("left", "<i4"), ("right", "<i4"), ("candidates", "<i4"),
                  ("iterations_run", "<i4"), ("converged", "<i4"),
                  ("final_max_delta", "<f4"), ("oscillating", "<i4")])


@dataclass
class MatchConfig:
    max_dy: float = 2.0
    min_disparity: float = 1.0
    max_disparity: float = 220.0
    max_candidates: int = 24
    sigma_y: float = 1.0
    w_y: float = 1.0
    w_prior: float = 1.0
    w_smooth: float = 1.0
    iterations: int = 20
    damping: float = 0.4
    lam: float = 0.0           # lambda: clutter, leaving a left keypoint unmatched
    gamma: float = 0.0         # misdetection, leaving a right keypoint unmatched
    converge_eps: float = 1e-4

    def to_c(self) -> DeMatchCfgC:
        c = DeMatchCfgC()
        for f in fields(self):
            setattr(c, "lambda_" if f.name == "lam" else f.name, getattr(self, f.name))
        return c


@dataclass
class MatchStats:
    """``doubleeye::MatchStats`` (match.hpp:178-186)."""
    left: int = 0
    right: int = 0
    candidates: int = 0
    iterations_run: int = 0
    converged: bool = False
    final_max_delta: float = 0.0
    oscillating: int = 0

    def _fill(self, st: DeMatchStatsC):
        for name, _ in DeMatchStatsC._fields_:
            v = getattr(st, name)
            setattr(self, name, bool(v) if name == "converged" else v)


def keypoints(xy) -> np.ndarray:
    """A ``KEYPOINT`` array from an (N, 2) array of x, y (or pass one through)."""
    a = np.asarray(xy)
    if a.dtype == KEYPOINT:
        return np.ascontiguousarray(a)
    a = np.asarray(a, np.float32).reshape(-1, 2)
    out = np.zeros(len(a), KEYPOINT)
    out["x"], out["y"] = a[:, 0], a[:, 1]
    return out


class _Keep:
    def __init__(self):
        self.refs = []

    def ptr(self, a, dtype):
        a = np.ascontiguousarray(a, dtype)
        self.refs.append(a)
        return a.ctypes.data_as(C.c_void_p) if a.size else None


def _priors(keep: _Keep, prior, smooth):
    pp = sp = None
    if prior is not None:
        p = DePriorC(int(prior["cell"]), int(prior["cols"]), int(prior["rows"]),
                     keep.ptr(prior["disparity"], np.float32), keep.ptr(prior["sigma"], np.float32),
                     keep.ptr(prior["valid"], np.uint8))
        keep.refs.append(p)
        pp = C.byref(p)
    if smooth is not None:
        s = DeSmoothC(len(smooth["disparity"]), keep.ptr(smooth["disparity"], np.float32),
                      keep.ptr(smooth["sigma"], np.float32), keep.ptr(smooth["valid"], np.uint8))
        keep.refs.append(s)
        sp = C.byref(s)
    return pp, sp


def _ctx(ctx):
    if ctx is None:
        from .dense import default_context
        ctx = default_context()
    return ctx


def _common(keep, left, desc_left, right, desc_right):
    L, R = keypoints(left), keypoints(right)
    dl = np.ascontiguousarray(desc_left, np.uint64)
    dr = np.ascontiguousarray(desc_right, np.uint64)
    if len(dl) != len(L) or len(dr) != len(R):
        raise ValueError("descriptors must be parallel to their keypoints")
    return (L, R, [keep.ptr(L, KEYPOINT), keep.ptr(dl, np.uint64), len(L),
                   keep.ptr(R, KEYPOINT), keep.ptr(dr, np.uint64), len(R)])


def _run(fn, left, desc_left, right, desc_right, cfg, prior, smooth, extra, stats, ctx):
    keep = _Keep()
    L, R, args = _common(keep, left, desc_left, right, desc_right)
    cfg = cfg or MatchConfig()
    out = np.zeros(max(1, min(len(L), len(R))), MATCH)
    n = C.c_int(0)
    st = DeMatchStatsC()
    pp, sp = _priors(keep, prior, smooth)
    check(fn(_ctx(ctx).h, *args, C.byref(cfg.to_c()), pp, sp, *extra,
             out.ctypes.data_as(C.c_void_p), C.byref(n), C.byref(st)))
    if stats is not None:
        stats._fill(st)
    return out[: n.value].copy()


def match_masda(left, desc_left, right, desc_right, cfg: MatchConfig | None = None,
                stats: MatchStats | None = None, prior=None, smooth=None, ctx=None):
    """``doubleeye::match_masda``: MASDA max-sum matching of left to right
    keypoints. ``prior`` is a dict (cell, cols, rows, disparity, sigma, valid),
    ``smooth`` a dict (disparity, sigma, valid) per left keypoint."""
    return _run(lib().de_match_masda, left, desc_left, right, desc_right, cfg, prior, smooth,
                [], stats, ctx)


def match_mutual_nn(left, desc_left, right, desc_right, cfg: MatchConfig | None = None,
                    ratio: float = 0.85, stats: MatchStats | None = None, prior=None,
                    smooth=None, ctx=None):
    """``doubleeye::match_mutual_nn``: mutual nearest neighbour + ratio test."""
    return _run(lib().de_match_mutual_nn, left, desc_left, right, desc_right, cfg, prior,
                smooth, [C.c_float(ratio)], stats, ctx)


def match_candidates(left, desc_left, right, desc_right, cfg: MatchConfig | None = None,
                     prior=None, smooth=None, ctx=None):
    """The candidate graph in the solver's edge order: (ei, ej, es)."""
    keep = _Keep()
    L, R, args = _common(keep, left, desc_left, right, desc_right)
    cfg = cfg or MatchConfig()
    cap = max(1, len(L) * max(1, cfg.max_candidates))
    ei = np.zeros(cap, np.int32)
    ej = np.zeros(cap, np.int32)
    es = np.zeros(cap, np.float32)
    n = C.c_int(0)
    pp, sp = _priors(keep, prior, smooth)
    check(lib().de_match_candidates(_ctx(ctx).h, *args, C.byref(cfg.to_c()), pp, sp,
                                    ei.ctypes.data_as(C.c_void_p), ej.ctypes.data_as(C.c_void_p),
                                    es.ctypes.data_as(C.c_void_p), C.byref(n)))
    return ei[: n.value].copy(), ej[: n.value].copy(), es[: n.value].copy()


def matching_objective(matches: np.ndarray, cfg: MatchConfig, n_left: int,
                       n_right: int) -> float:
    """Sum of s(i, j) over the matching plus lambda per unmatched left keypoint
    and gamma per unmatched right one (match.cpp:812-819), in float32."""
    t = np.float32(0.0)
    for s in np.asarray(matches["score"], np.float32):
        t = np.float32(t + s)
    n = len(matches)
    t = np.float32(t + np.float32(cfg.lam) * np.float32(n_left - n))
    t = np.float32(t + np.float32(cfg.gamma) * np.float32(n_right - n))
    return float(t)


@dataclass
class SparseFrame:
    """Device buffers of one frame of a sparse batch (DeviceArray objects)."""
    left: object
    desc_left: object
    right: object
    desc_right: object
    out: object
    n_out: object
    stats: object

    def to_c(self) -> DeSparseFrameC:
        return DeSparseFrameC(self.left.ptr, self.desc_left.ptr, self.left.shape[0],
                              self.right.ptr, self.desc_right.ptr, self.right.shape[0],
                              self.out.ptr, self.n_out.ptr, self.stats.ptr)


def match_masda_batch_dev(ctx, frames, cfgs, stream: int = 0):
    """``match_masda`` over device-resident frames, one launch per stage."""
    n = len(frames)
    fa = (DeSparseFrameC * n)(*[f.to_c() for f in frames])
    ca = (DeMatchCfgC * n)(*[c.to_c() for c in cfgs])
    check(lib().de_match_masda_batch_dev(ctx.h, fa, ca, n, C.c_void_p(stream or None)))
