"""ctypes binding of the keypoint-graph BP oracle (oracle/knn_bp_ref.cpp in
oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py as the checker. The product package never imports
this module.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import binding

NLAB = 8
OUT_DTYPE = np.dtype([("d1", "<i4"), ("d2", "<i4"), ("c1", "<f4"), ("c2", "<f4")])


class OrcKnnCfg(C.Structure):
    _fields_ = [("census_hw", C.c_int), ("census_hh", C.c_int), ("dmax", C.c_int),
                ("n_cand", C.c_int), ("k_nn", C.c_int), ("iters", C.c_int),
                ("w_data", C.c_float), ("lambda_", C.c_float), ("trunc", C.c_int)]


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


def cfg(**kw) -> OrcKnnCfg:
    c = OrcKnnCfg()
    binding.lib().orc_kb_default_cfg(C.byref(c))
    for k, v in kw.items():
        setattr(c, "lambda_" if k in ("lam", "lambda") else k, v)
    return c


def candidates(L, R, kp, c: OrcKnnCfg):
    L = np.ascontiguousarray(L, np.uint8)
    R = np.ascontiguousarray(R, np.uint8)
    kp = np.ascontiguousarray(kp, np.int32).reshape(-1, 2)
    n = len(kp)
    cd = np.zeros((n, NLAB), np.int32)
    cc = np.zeros((n, NLAB), np.float32)
    nc = np.zeros(n, np.int32)
    rc = binding.lib().orc_kb_candidates(_p(L), _p(R), L.shape[1], L.shape[0], _p(kp), n,
                                         C.byref(c), _p(cd), _p(cc), _p(nc))
    if rc:
        raise ValueError("keypoint census window outside the image")
    return cd, cc, nc


def knn_graph(kp, k: int):
    kp = np.ascontiguousarray(kp, np.int32).reshape(-1, 2)
    n = len(kp)
    cap = n * (2 * k + 1) + 1
    off = np.zeros(n + 1, np.int32)
    adj = np.zeros(cap, np.int32)
    e = binding.lib().orc_kb_knn_graph(_p(kp), n, k, _p(off), _p(adj), cap)
    if e < 0:
        raise ValueError("adjacency capacity")
    return off, adj[:e].copy()


def bp_graph(cd, cc, nc, off, adj, c: OrcKnnCfg, want_beliefs=False):
    cd = np.ascontiguousarray(cd, np.int32)
    cc = np.ascontiguousarray(cc, np.float32)
    nc = np.ascontiguousarray(nc, np.int32)
    off = np.ascontiguousarray(off, np.int32)
    adj = np.ascontiguousarray(adj, np.int32)
    n = len(nc)
    out = np.zeros(n, OUT_DTYPE)
    bel = np.zeros((n, NLAB), np.float32) if want_beliefs else None
    rc = binding.lib().orc_kb_bp_graph(n, _p(cd), _p(cc), _p(nc), _p(off), _p(adj), C.byref(c),
                                       _p(out), _p(bel) if bel is not None else None)
    if rc:
        raise ValueError("graph is not symmetric")
    return (out, bel) if want_beliefs else out


def run(L, R, kp, c: OrcKnnCfg):
    L = np.ascontiguousarray(L, np.uint8)
    R = np.ascontiguousarray(R, np.uint8)
    kp = np.ascontiguousarray(kp, np.int32).reshape(-1, 2)
    out = np.zeros(len(kp), OUT_DTYPE)
    rc = binding.lib().orc_kb_run(_p(L), _p(R), L.shape[1], L.shape[0], _p(kp), len(kp),
                                  C.byref(c), _p(out))
    if rc:
        raise ValueError("orc_kb_run failed")
    return out


def brute_min_marginals(cd, cc, nc, off, adj, c: OrcKnnCfg) -> np.ndarray:
    cd = np.ascontiguousarray(cd, np.int32)
    cc = np.ascontiguousarray(cc, np.float32)
    nc = np.ascontiguousarray(nc, np.int32)
    off = np.ascontiguousarray(off, np.int32)
    adj = np.ascontiguousarray(adj, np.int32)
    mm = np.zeros((len(nc), NLAB), np.float64)
    rc = binding.lib().orc_kb_brute_min_marginals(len(nc), _p(cd), _p(cc), _p(nc), _p(off),
                                                  _p(adj), C.byref(c), _p(mm))
    if rc:
        raise ValueError("every node needs a label")
    return mm
