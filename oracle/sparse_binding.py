"""ctypes binding of the sparse-matcher oracle (oracle/liboracle.so) and of the
reference harness built from the reference's own match.cpp
(oracle/_ref/libref_sparse.so, present only where /root/reference was).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py as the checker. The product package never imports
this module.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from . import binding

HERE = os.path.dirname(os.path.abspath(__file__))
REF_SPARSE = os.path.join(HERE, "_ref", "libref_sparse.so")

KP_DTYPE = np.dtype([("x", "<f4"), ("y", "<f4"), ("response", "<f4"), ("local_std", "<f4")])
MATCH_DTYPE = np.dtype([("left", "<i4"), ("right", "<i4"), ("score", "<f4"), ("belief", "<f4"),
                        ("disparity", "<f4"), ("dy", "<f4"), ("margin", "<f4")])


class OrcMatchCfg(C.Structure):
    _fields_ = [("max_dy", C.c_float), ("min_disparity", C.c_float),
                ("max_disparity", C.c_float), ("max_candidates", C.c_int),
                ("sigma_y", C.c_float), ("w_y", C.c_float), ("w_prior", C.c_float),
                ("w_smooth", C.c_float), ("iterations", C.c_int), ("damping", C.c_float),
                ("lambda_", C.c_float), ("gamma", C.c_float), ("converge_eps", C.c_float)]


class OrcPrior(C.Structure):
    _fields_ = [("cell", C.c_int), ("cols", C.c_int), ("rows", C.c_int),
                ("disparity", C.c_void_p), ("sigma", C.c_void_p), ("valid", C.c_void_p)]


class OrcSmooth(C.Structure):
    _fields_ = [("n", C.c_int), ("disparity", C.c_void_p), ("sigma", C.c_void_p),
                ("valid", C.c_void_p)]


class OrcMatchStats(C.Structure):
    _fields_ = [("left", C.c_int), ("right", C.c_int), ("candidates", C.c_int),
                ("iterations_run", C.c_int), ("converged", C.c_int),
                ("final_max_delta", C.c_float), ("oscillating", C.c_int)]


def match_cfg(**kw) -> OrcMatchCfg:
    c = OrcMatchCfg()
    binding.lib().orc_match_default_cfg(C.byref(c))
    for k, v in kw.items():
        setattr(c, "lambda_" if k in ("lam", "lambda") else k, v)
    return c


def keypoints(xy) -> np.ndarray:
    """Structured keypoint array from an (N, 2) x/y array (or pass-through)."""
    a = np.asarray(xy)
    if a.dtype == KP_DTYPE:
        return np.ascontiguousarray(a)
    a = np.asarray(a, np.float32).reshape(-1, 2)
    out = np.zeros(len(a), KP_DTYPE)
    out["x"], out["y"] = a[:, 0], a[:, 1]
    out["response"] = 10.0
    out["local_std"] = 10.0
    return out


class _Holder:
    """Keeps numpy buffers alive while ctypes structs point into them."""

    def __init__(self):
        self.keep = []

    def arr(self, a, dt):
        a = np.ascontiguousarray(a, dt)
        self.keep.append(a)
        return a.ctypes.data_as(C.c_void_p)


def _prior(prior, h: _Holder):
    if prior is None:
        return None
    p = OrcPrior(int(prior["cell"]), int(prior["cols"]), int(prior["rows"]),
                 h.arr(prior["disparity"], np.float32), h.arr(prior["sigma"], np.float32),
                 h.arr(prior["valid"], np.uint8))
    h.keep.append(p)
    return C.byref(p)


def _smooth(sm, h: _Holder):
    if sm is None:
        return None
    s = OrcSmooth(len(sm["disparity"]), h.arr(sm["disparity"], np.float32),
                  h.arr(sm["sigma"], np.float32), h.arr(sm["valid"], np.uint8))
    h.keep.append(s)
    return C.byref(s)


def _stats_dict(st: OrcMatchStats) -> dict:
    return {k: getattr(st, k) for k, _ in OrcMatchStats._fields_}


_ref = None


def ref_lib():
    global _ref
    if _ref is None:
        if not os.path.exists(REF_SPARSE):
            raise FileNotFoundError(f"{REF_SPARSE} not built (make -f oracle/Makefile.ref)")
        _ref = C.CDLL(REF_SPARSE)
        _ref.ref_matching_objective.restype = C.c_float
    return _ref


def _run(fn, L, dl, R, dr, c, extra_pre, extra_post, prior=None, smooth=None, stats=True):
    h = _Holder()
    L, R = keypoints(L), keypoints(R)
    nl, nr = len(L), len(R)
    out = np.zeros(max(1, min(nl, nr)), MATCH_DTYPE)
    n = C.c_int(0)
    st = OrcMatchStats()
    args = [h.arr(L, KP_DTYPE), h.arr(np.asarray(dl, np.uint64), np.uint64), nl,
            h.arr(R, KP_DTYPE), h.arr(np.asarray(dr, np.uint64), np.uint64), nr, C.byref(c)]
    args += extra_pre(h, prior, smooth)
    args += extra_post + [out.ctypes.data_as(C.c_void_p), C.byref(n)]
    if stats:
        args.append(C.byref(st))
    rc = fn(*args)
    if rc:
        raise ValueError(f"{fn.__name__} rc={rc}")
    return out[: n.value].copy(), (_stats_dict(st) if stats else None)


def _ps(h, prior, smooth):
    return [_prior(prior, h), _smooth(smooth, h)]


def match_masda(L, dl, R, dr, c: OrcMatchCfg, prior=None, smooth=None, canonical=True):
    """Oracle match_masda. Returns (matches structured array, stats dict)."""
    return _run(binding.lib().orc_match_masda, L, dl, R, dr, c, _ps,
                [int(canonical)], prior, smooth)


def match_mutual_nn(L, dl, R, dr, c: OrcMatchCfg, ratio=0.85, prior=None, smooth=None,
                    canonical=True):
    return _run(binding.lib().orc_match_mutual_nn, L, dl, R, dr, c, _ps,
                [C.c_float(ratio), int(canonical)], prior, smooth)


def match_brute_force(L, dl, R, dr, c: OrcMatchCfg):
    return _run(binding.lib().orc_match_brute_force, L, dl, R, dr, c,
                lambda h, p, s: [], [], stats=False)[0]


def candidates(L, dl, R, dr, c: OrcMatchCfg, prior=None, smooth=None, canonical=True):
    """Candidate edges (ei, ej, es) in the solver's edge order."""
    h = _Holder()
    L, R = keypoints(L), keypoints(R)
    cap = len(L) * max(1, c.max_candidates)
    ei = np.zeros(max(cap, 1), np.int32)
    ej = np.zeros(max(cap, 1), np.int32)
    es = np.zeros(max(cap, 1), np.float32)
    n = binding.lib().orc_match_candidates(
        h.arr(L, KP_DTYPE), h.arr(np.asarray(dl, np.uint64), np.uint64), len(L),
        h.arr(R, KP_DTYPE), h.arr(np.asarray(dr, np.uint64), np.uint64), len(R), C.byref(c),
        _prior(prior, h), _smooth(smooth, h), int(canonical),
        ei.ctypes.data_as(C.c_void_p), ej.ctypes.data_as(C.c_void_p),
        es.ctypes.data_as(C.c_void_p))
    return ei[:n].copy(), ej[:n].copy(), es[:n].copy()


def objective(m: np.ndarray, c: OrcMatchCfg, n_left: int, n_right: int) -> float:
    """matching_objective (match.cpp:812-819), float32 accumulation in order."""
    t = np.float32(0.0)
    for s in m["score"]:
        t = np.float32(t + np.float32(s))
    t = np.float32(t + np.float32(c.lambda_) * np.float32(n_left - len(m)))
    t = np.float32(t + np.float32(c.gamma) * np.float32(n_right - len(m)))
    return float(t)


# --- the reference itself -------------------------------------------------------

def ref_match_masda(L, dl, R, dr, c: OrcMatchCfg, prior=None, smooth=None):
    return _run(ref_lib().ref_match_masda, L, dl, R, dr, c, _ps, [], prior, smooth)


def ref_match_mutual_nn(L, dl, R, dr, c: OrcMatchCfg, ratio=0.85, prior=None, smooth=None):
    return _run(ref_lib().ref_match_mutual_nn, L, dl, R, dr, c, _ps, [C.c_float(ratio)],
                prior, smooth)


def ref_match_brute_force(L, dl, R, dr, c: OrcMatchCfg):
    return _run(ref_lib().ref_match_brute_force, L, dl, R, dr, c,
                lambda h, p, s: [], [], stats=False)[0]
