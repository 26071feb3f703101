"""The edge-list MASDA functions of the reference's article/masda_stereo.py
and article/dense_masda.py, on the MI355X through the C ABI (float64 messages,
numpy's rounding).

Same names, arguments and return types:

* ``masda_sparse(ei, ej, se, m, n, lam, gam, iters, damping, eps)`` ->
  ``(dict left -> right, per-iteration change history)`` (masda_stereo.py:453)
* ``masda(S, lam, gam, iters, damping, eps)`` -> same, for a dense score
  matrix with -inf where there is no candidate (masda_stereo.py:329)
* ``decide(S, belief, lam)`` -> dict (masda_stereo.py:392)
* ``_seg_max_excluding(vals, idx, n)`` (masda_stereo.py:417)
* ``dense_problem(left, right, dmin, dmax, census_cfg)`` ->
  ``(ei, ej, se, n_nodes)`` (dense_masda.py:39)
* ``dense_masda(left, right, dmin, dmax, ...)`` -> disparity map, the device
  pipeline of dense_masda.run (dense_masda.py:82) without its margin gate.

The returned dicts hold the same pairs as the reference's; they are built in
ascending left index, where the reference inserts in belief order. Exactly
equal beliefs are taken in edge-index order (the reference: numpy's unstable
argsort).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from ._lib import check, lib

__all__ = ["masda_sparse", "masda", "decide", "_seg_max_excluding", "dense_problem",
           "dense_masda"]


def _ctx(ctx):
    if ctx is None:
        from .dense import default_context
        ctx = default_context()
    return ctx


def _p(a):
    return a.ctypes.data_as(C.c_void_p) if a is not None and a.size else None


def _dict(assign: np.ndarray) -> dict:
    idx = np.nonzero(assign >= 0)[0]
    return {int(i): int(assign[i]) for i in idx}


def masda_sparse(ei, ej, se, m, n, lam=-0.1, gam=-0.1, iters=30, damping=0.4, eps=1e-5,
                 ctx=None):
    ei = np.ascontiguousarray(ei, np.int64)
    ej = np.ascontiguousarray(ej, np.int64)
    se = np.asarray(se)
    if se.dtype == np.float32:
        fn, se = lib().de_masda_edges_f32, np.ascontiguousarray(se)
    else:
        fn, se = lib().de_masda_edges, np.ascontiguousarray(se, np.float64)
    if not (len(ei) == len(ej) == len(se)):
        raise ValueError("ei, ej and se must have one entry per edge")
    assign = np.full(max(int(m), 1), -1, np.int64)
    hist = np.zeros(max(int(iters), 1), np.float64)
    nit = C.c_int(0)
    check(fn(_ctx(ctx).h, _p(ei), _p(ej), _p(se), len(se), int(m), int(n), float(lam),
             float(gam), int(iters), float(damping), float(eps), _p(assign), _p(hist),
             C.byref(nit)))
    return _dict(assign[: int(m)]), hist[: nit.value].copy()


def masda(S, lam=-0.1, gam=-0.1, iters=30, damping=0.4, eps=1e-5, ctx=None):
    S = np.ascontiguousarray(S, np.float32)
    m, n = S.shape
    assign = np.full(max(m, 1), -1, np.int64)
    hist = np.zeros(max(int(iters), 1), np.float64)
    nit = C.c_int(0)
    check(lib().de_masda_matrix(_ctx(ctx).h, _p(S), m, n, float(lam), float(gam), int(iters),
                                float(damping), float(eps), _p(assign), _p(hist), C.byref(nit)))
    return _dict(assign[:m]), hist[: nit.value].copy()


def decide(S, belief, lam, ctx=None):
    S = np.asarray(S, np.float32)
    m, n = S.shape
    bel = np.asarray(belief, np.float64)
    # edges: every entry of the matrix, row-major (non-finite beliefs are never taken)
    ei, ej = np.divmod(np.arange(m * n, dtype=np.int64), n)
    se = np.ascontiguousarray(S.ravel(), np.float64)
    bf = np.ascontiguousarray(bel.ravel())
    assign = np.full(max(m, 1), -1, np.int64)
    check(lib().de_masda_decide(_ctx(ctx).h, _p(ei), _p(ej), _p(se), _p(bf), m * n, m, n,
                                float(lam), _p(assign)))
    return _dict(assign[:m])


def _seg_max_excluding(vals, idx, n, ctx=None):
    vals = np.ascontiguousarray(vals, np.float64)
    idx = np.ascontiguousarray(idx, np.int64)
    out = np.empty(len(vals), np.float64)
    check(lib().de_seg_max_excluding(_ctx(ctx).h, _p(vals), _p(idx), len(vals), int(n), _p(out)))
    return out


def dense_problem(left, right, dmin, dmax, census_cfg=(3, 3), ctx=None):
    L = np.ascontiguousarray(left, np.uint8)
    R = np.ascontiguousarray(right, np.uint8)
    H, W = L.shape
    hw, hh = census_cfg
    c = _ctx(ctx)
    E = C.c_int64(0)
    check(lib().de_dense_problem(c.h, _p(L), _p(R), W, H, int(dmin), int(dmax), hw, hh, None,
                                 None, None, 0, C.byref(E)))
    ei = np.empty(E.value, np.int64)
    ej = np.empty(E.value, np.int64)
    se = np.empty(E.value, np.float32)
    check(lib().de_dense_problem(c.h, _p(L), _p(R), W, H, int(dmin), int(dmax), hw, hh, _p(ei),
                                 _p(ej), _p(se), E.value, C.byref(E)))
    return ei, ej, se, H * W


def dense_masda(left, right, dmin, dmax, lam=-0.1, gam=-0.1, iters=12, damping=0.4, eps=1e-5,
                census_cfg=(3, 3), ctx=None):
    """Disparity map (NaN where unmatched) and the change history."""
    L = np.ascontiguousarray(left, np.uint8)
    R = np.ascontiguousarray(right, np.uint8)
    H, W = L.shape
    disp = np.empty((H, W), np.float32)
    hist = np.zeros(max(int(iters), 1), np.float64)
    nit = C.c_int(0)
    check(lib().de_dense_masda(_ctx(ctx).h, _p(L), _p(R), W, H, int(dmin), int(dmax),
                               census_cfg[0], census_cfg[1], float(lam), float(gam), int(iters),
                               float(damping), float(eps), _p(disp), _p(hist), C.byref(nit)))
    return disp, hist[: nit.value].copy()
