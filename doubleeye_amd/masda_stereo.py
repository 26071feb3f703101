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
* ``census(img, pts, half_w, half_h)`` -> uint64 descriptor per point
  (masda_stereo.py:264)
* ``hamming(a, b)`` -> int32 popcount of a ^ b, broadcast (masda_stereo.py:289)
* ``build_problem(pl, dl, pr, dr, max_dy, dmin, dmax, sigma_y, w_y)`` ->
  float32 m x n score matrix, -inf where no candidate (masda_stereo.py:305),
  with the reference's 48-bit scaling (``BITS``).

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
           "dense_masda", "census", "hamming", "build_problem", "BITS"]

BITS = 48   # descriptor length the score scale assumes (masda_stereo.py:302)


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


def census(img, pts, half_w=3, half_h=3, ctx=None):
    """Census descriptor at each point: bit b in scan order (dy-major, dx-minor,
    centre skipped) set when that neighbour is darker than the centre. Points
    are (x, y), truncated to integers as ``pts.astype(int)`` does; indexing
    follows numpy's (a negative index counts from the end; past the end raises
    IndexError)."""
    img = np.ascontiguousarray(img, np.uint8)
    H, W = img.shape
    xy = np.ascontiguousarray(np.asarray(pts).astype(int), np.int64).reshape(-1, 2)
    n = len(xy)
    out = np.zeros(n, np.uint64)
    if n == 0:
        return out
    hw, hh = int(half_w), int(half_h)
    bad = ((xy[:, 0] - hw < -W) | (xy[:, 0] + hw >= W) | (xy[:, 1] - hh < -H) |
           (xy[:, 1] + hh >= H))
    if bad.any():
        i = int(np.nonzero(bad)[0][0])
        raise IndexError(f"census window of point {i} {tuple(xy[i])} is outside the "
                         f"{H}x{W} image")
    check(lib().de_census_points(_ctx(ctx).h, _p(img), W, H, _p(xy), n, hw, hh, _p(out)))
    return out


def hamming(a, b, ctx=None):
    """popcount(a ^ b) elementwise over the broadcast shape, int32."""
    a, b = np.broadcast_arrays(np.asarray(a, np.uint64), np.asarray(b, np.uint64))
    shape = a.shape
    a = np.ascontiguousarray(a).ravel()
    b = np.ascontiguousarray(b).ravel()
    out = np.zeros(a.size, np.int32)
    if a.size:
        check(lib().de_hamming(_ctx(ctx).h, _p(a), _p(b), a.size, _p(out)))
    return out.reshape(shape)


def build_problem(pl, dl, pr, dr, max_dy=1.5, dmin=1.0, dmax=60.0, sigma_y=1.0, w_y=1.0,
                  ctx=None):
    """Sparse candidate graph as a score matrix: S[i, j] = (BITS/2 - h)/(BITS/2)
    - w_y (dy/sigma_y)^2 for |dy| <= max_dy and dmin <= d <= dmax, -inf
    elsewhere (float32, m x n). float32 points are evaluated in numpy's float32
    arithmetic, others in float64."""
    pl = np.asarray(pl)
    pr = np.asarray(pr)
    fp64 = not (pl.dtype == np.float32 and pr.dtype == np.float32)
    t = np.float64 if fp64 else np.float32
    pl = np.ascontiguousarray(pl, t).reshape(-1, 2)
    pr = np.ascontiguousarray(pr, t).reshape(-1, 2)
    dl = np.ascontiguousarray(dl, np.uint64).ravel()
    dr = np.ascontiguousarray(dr, np.uint64).ravel()
    m, n = len(pl), len(pr)
    if len(dl) != m or len(dr) != n:
        raise ValueError("one descriptor per point")
    S = np.full((m, n), -np.inf, np.float32)
    if m and n:
        check(lib().de_build_problem(_ctx(ctx).h, _p(pl), _p(dl), m, _p(pr), _p(dr), n,
                                     int(fp64), float(max_dy), float(dmin), float(dmax),
                                     float(sigma_y), float(w_y), _p(S)))
    return S
