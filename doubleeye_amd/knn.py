"""Keypoint-graph belief propagation on the MI355X through the C ABI.

Labels along the epipolar row by census Hamming distance, a symmetric k-NN
graph over the keypoints, min-sum loopy BP with a truncated-linear
disparity-difference term, top-2 labels per keypoint (include/doubleeye_hip.h,
"Keypoint-graph belief propagation"). The reference has no such solver; this
module is its only interface.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, fields

import numpy as np

from ._lib import DeKnnCfgC, DeKnnFrameC, check, lib

__all__ = ["KnnConfig", "OUT", "knn_bp", "knn_candidates", "knn_graph", "knn_bp_graph",
           "KnnFrame", "knn_bp_batch_dev"]

NLAB = 8
OUT = np.dtype([("d1", "<i4"), ("d2", "<i4"), ("c1", "<f4"), ("c2", "<f4")])


@dataclass
class KnnConfig:
    census_hw: int = 4
    census_hh: int = 3
    dmax: int = 127
    n_cand: int = 8
    k_nn: int = 8
    iters: int = 10
    w_data: float = 1.0
    lam: float = 1.0           # smoothness cost per pixel of disparity difference
    trunc: int = 6

    def to_c(self) -> DeKnnCfgC:
        c = DeKnnCfgC()
        for f in fields(self):
            setattr(c, "lambda_" if f.name == "lam" else f.name, getattr(self, f.name))
        return c


def _p(a):
    return a.ctypes.data_as(C.c_void_p) if a is not None and a.size else None


def _ctx(ctx):
    if ctx is None:
        from .dense import default_context
        ctx = default_context()
    return ctx


def _inputs(L, R, kp):
    L = np.ascontiguousarray(L, np.uint8)
    R = np.ascontiguousarray(R, np.uint8)
    kp = np.ascontiguousarray(kp, np.int32).reshape(-1, 2)
    if L.shape != R.shape:
        raise ValueError("left and right images differ in shape")
    return L, R, kp


def knn_bp(L, R, kp, cfg: KnnConfig | None = None, ctx=None) -> np.ndarray:
    """The whole workload on one frame; returns an OUT array per keypoint."""
    L, R, kp = _inputs(L, R, kp)
    out = np.zeros(len(kp), OUT)
    check(lib().de_knn_bp(_ctx(ctx).h, _p(L), _p(R), L.shape[1], L.shape[0], _p(kp), len(kp),
                          C.byref((cfg or KnnConfig()).to_c()), _p(out)))
    return out


def knn_candidates(L, R, kp, cfg: KnnConfig | None = None, ctx=None):
    """Labels: (cand_d [n][8], cand_c [n][8], ncand [n])."""
    L, R, kp = _inputs(L, R, kp)
    n = len(kp)
    cd = np.zeros((n, NLAB), np.int32)
    cc = np.zeros((n, NLAB), np.float32)
    nc = np.zeros(n, np.int32)
    check(lib().de_knn_candidates(_ctx(ctx).h, _p(L), _p(R), L.shape[1], L.shape[0], _p(kp), n,
                                  C.byref((cfg or KnnConfig()).to_c()), _p(cd), _p(cc), _p(nc)))
    return cd, cc, nc


def knn_graph(L, R, kp, cfg: KnnConfig | None = None, ctx=None):
    """The k-NN graph as CSR: (off [n+1], adj)."""
    L, R, kp = _inputs(L, R, kp)
    n = len(kp)
    off = np.zeros(n + 1, np.int32)
    cap = max(1, n * 64)
    adj = np.zeros(cap, np.int32)
    ne = C.c_int(0)
    check(lib().de_knn_graph(_ctx(ctx).h, _p(L), _p(R), L.shape[1], L.shape[0], _p(kp), n,
                             C.byref((cfg or KnnConfig()).to_c()), _p(off), _p(adj), cap,
                             C.byref(ne)))
    return off, adj[: ne.value].copy()


def knn_bp_graph(cand_d, cand_c, ncand, off, adj, cfg: KnnConfig | None = None,
                 want_beliefs: bool = False, ctx=None):
    """BP alone on a given symmetric graph (ascending CSR lists)."""
    cd = np.ascontiguousarray(cand_d, np.int32).reshape(-1, NLAB)
    cc = np.ascontiguousarray(cand_c, np.float32).reshape(-1, NLAB)
    nc = np.ascontiguousarray(ncand, np.int32)
    off = np.ascontiguousarray(off, np.int32)
    adj = np.ascontiguousarray(adj, np.int32)
    n = len(nc)
    out = np.zeros(n, OUT)
    bel = np.zeros((n, NLAB), np.float32) if want_beliefs else None
    check(lib().de_knn_bp_graph(_ctx(ctx).h, n, _p(cd), _p(cc), _p(nc), _p(off), _p(adj),
                                C.byref((cfg or KnnConfig()).to_c()), _p(out), _p(bel)))
    return (out, bel) if want_beliefs else out


@dataclass
class KnnFrame:
    """Device buffers of one frame of a batch (DeviceArray objects)."""
    L: object
    R: object
    kp: object
    out: object

    def to_c(self) -> DeKnnFrameC:
        H, W = self.L.shape
        return DeKnnFrameC(self.L.ptr, self.R.ptr, W, H, self.kp.ptr, self.kp.shape[0],
                           self.out.ptr)


def knn_bp_batch_dev(ctx, frames, cfgs, stream: int = 0):
    n = len(frames)
    fa = (DeKnnFrameC * n)(*[f.to_c() for f in frames])
    ca = (DeKnnCfgC * n)(*[c.to_c() for c in cfgs])
    check(lib().de_knn_bp_batch_dev(ctx.h, fa, ca, n, C.c_void_p(stream or None)))
