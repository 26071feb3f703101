"""ctypes binding of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py as the checker. The product package never imports
this module.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
LIB_PATH = os.path.join(HERE, "liboracle.so")


class OrcCfg(C.Structure):
    _fields_ = [("dmin", C.c_int), ("dmax", C.c_int), ("iters", C.c_int),
                ("census_hw", C.c_int), ("census_hh", C.c_int),
                ("sigma_s", C.c_float), ("sigma_r", C.c_float),
                ("ad", C.c_float), ("ad_trunc", C.c_int),
                ("lambda_", C.c_float), ("gamma", C.c_float),
                ("damping", C.c_float), ("min_margin", C.c_float),
                ("min_ratio", C.c_float),
                ("subpixel", C.c_int), ("fit_eq", C.c_int), ("lrc", C.c_int)]


def build(force: bool = False) -> str:
    """Compile the oracle with oracle/Makefile (g++), return the .so path."""
    cmd = ["make", "-s", "-f", "oracle/Makefile"] + (["-B"] if force else [])
    subprocess.run(cmd, cwd=ROOT, check=True)
    return LIB_PATH


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        P = C.c_void_p
        L.orc_default_cfg.argtypes = [C.POINTER(OrcCfg)]
        L.orc_census.argtypes = [P, C.c_int, C.c_int, C.c_int, C.c_int, P]
        L.orc_rf_table.argtypes = [C.c_float, C.c_float, P]
        L.orc_rf_coeffs.argtypes = [P, C.c_int, C.c_int, C.c_float, C.c_float, P, P]
        L.orc_score_tables.argtypes = [C.c_int, C.c_float, C.c_int, P, P, P]
        L.orc_filtered_planes_i16.argtypes = [P, P, C.c_int, C.c_int, C.POINTER(OrcCfg),
                                              C.c_int, C.c_int, P]
        L.orc_dense_run.argtypes = [P, P, C.c_int, C.c_int, C.POINTER(OrcCfg), P, P, P, P]
        L.orc_volume_f32.argtypes = [P, P, C.c_int, C.c_int, C.POINTER(OrcCfg), P]
        L.orc_dense_run_full.argtypes = [P, P, C.c_int, C.c_int, C.POINTER(OrcCfg), P, P, P]
        L.orc_confidence.argtypes = [C.c_float] * 5
        L.orc_confidence.restype = C.c_float
        L.orc_confidence_u8.argtypes = [C.c_float] * 5
        L.orc_confidence_u8.restype = C.c_uint8
        L.orc_dedd_planes.argtypes = [P, P, P, P, C.c_int, C.c_float, P, P, P]
        for name in ("orc_sparse_match_masda", "orc_masda_edges", "orc_knn_bp"):
            if hasattr(L, name):
                getattr(L, name).restype = C.c_int
        _lib = L
    return _lib


def _p(a: np.ndarray):
    return a.ctypes.data_as(C.c_void_p)


def cfg(**kw) -> OrcCfg:
    c = OrcCfg()
    lib().orc_default_cfg(C.byref(c))
    for k, v in kw.items():
        setattr(c, "lambda_" if k == "lam" else k, v)
    return c


def census(img: np.ndarray, hw: int = 3, hh: int = 3) -> np.ndarray:
    img = np.ascontiguousarray(img, np.uint8)
    out = np.empty(img.shape, np.uint64)
    rc = lib().orc_census(_p(img), img.shape[1], img.shape[0], hw, hh, _p(out))
    if rc:
        raise ValueError(f"orc_census rc={rc}")
    return out


def rf_table(sigma_s: float, sigma_r: float) -> np.ndarray:
    t = np.empty(256, np.uint16)
    lib().orc_rf_table(sigma_s, sigma_r, _p(t))
    return t


def rf_coeffs(L: np.ndarray, sigma_s: float, sigma_r: float):
    L = np.ascontiguousarray(L, np.uint8)
    ax = np.empty(L.shape, np.uint16)
    ay = np.empty(L.shape, np.uint16)
    lib().orc_rf_coeffs(_p(L), L.shape[1], L.shape[0], sigma_s, sigma_r, _p(ax), _p(ay))
    return ax, ay


def score_tables(half: int, ad: float, ad_trunc: int):
    tbl = np.empty(65, np.int16)
    adt = np.empty(256, np.int16)
    wq = np.zeros(1, np.int32)
    lib().orc_score_tables(half, ad, ad_trunc, _p(tbl), _p(adt), _p(wq))
    return tbl, adt, int(wq[0])


def filtered_planes(L, R, c: OrcCfg, k0: int, k1: int) -> np.ndarray:
    L = np.ascontiguousarray(L, np.uint8)
    R = np.ascontiguousarray(R, np.uint8)
    H, W = L.shape
    out = np.empty((k1 - k0, H, W), np.int16)
    rc = lib().orc_filtered_planes_i16(_p(L), _p(R), W, H, C.byref(c), k0, k1, _p(out))
    if rc:
        raise ValueError(f"orc_filtered_planes_i16 rc={rc}")
    return out


def dense_run(L, R, c: OrcCfg, want_cand: bool = False, want_ties: bool = False):
    """Returns (disp, margin, cand6 or None, tie_mask or None)."""
    L = np.ascontiguousarray(L, np.uint8)
    R = np.ascontiguousarray(R, np.uint8)
    H, W = L.shape
    disp = np.empty((H, W), np.float32)
    margin = np.empty((H, W), np.float32)
    cand = np.empty((6, H, W), np.float32) if want_cand else None
    ties = np.zeros((H, W), np.uint8) if want_ties else None
    rc = lib().orc_dense_run(_p(L), _p(R), W, H, C.byref(c), _p(disp), _p(margin),
                             _p(cand) if cand is not None else None,
                             _p(ties) if ties is not None else None)
    if rc:
        raise ValueError(f"orc_dense_run rc={rc}")
    return disp, margin, cand, ties


def dense_run_prior(L, R, c: OrcCfg, prior, band: int = 2, want_cand: bool = False,
                    want_ties: bool = False):
    """--prior: (disp, margin, cand6 or None, tie_mask or None). `prior` is
    W*H, or W/2 x H/2 (nearest-upsampled, doubled)."""
    L = np.ascontiguousarray(L, np.uint8)
    R = np.ascontiguousarray(R, np.uint8)
    H, W = L.shape
    pr = np.ascontiguousarray(prior, np.float32)
    if pr.size == (W // 2) * (H // 2) and pr.size != W * H:
        full = np.empty((H, W), np.float32)
        lib().orc_prior_upsample(_p(pr), W, H, _p(full))
        pr = full
    pr = pr.reshape(H, W)
    disp = np.empty((H, W), np.float32)
    margin = np.empty((H, W), np.float32)
    cand = np.empty((6, H, W), np.float32) if want_cand else None
    ties = np.zeros((H, W), np.uint8) if want_ties else None
    rc = lib().orc_dense_run_prior(_p(L), _p(R), W, H, C.byref(c), _p(pr), int(band), _p(disp),
                                   _p(margin), _p(cand) if cand is not None else None,
                                   _p(ties) if ties is not None else None)
    if rc:
        raise ValueError(f"orc_dense_run_prior rc={rc}")
    return disp, margin, cand, ties


def dense_run_c2f(L, R, c: OrcCfg, band: int = 2, want_cand: bool = False,
                  want_ties: bool = False):
    """--c2f: (disp, margin, cand6 or None, tie_mask or None, coarse, used_mask)."""
    L = np.ascontiguousarray(L, np.uint8)
    R = np.ascontiguousarray(R, np.uint8)
    H, W = L.shape
    disp = np.empty((H, W), np.float32)
    margin = np.empty((H, W), np.float32)
    cand = np.empty((6, H, W), np.float32) if want_cand else None
    ties = np.zeros((H, W), np.uint8) if want_ties else None
    coarse = np.empty((H // 2, W // 2), np.float32)
    used = C.c_int(0)
    rc = lib().orc_dense_run_c2f(_p(L), _p(R), W, H, C.byref(c), int(band), _p(disp),
                                 _p(margin), _p(cand) if cand is not None else None,
                                 _p(ties) if ties is not None else None, _p(coarse),
                                 C.byref(used))
    if rc:
        raise ValueError(f"orc_dense_run_c2f rc={rc}")
    return disp, margin, cand, ties, coarse, bool(used.value)


def dense_run_full(L, R, c: OrcCfg, want_ties: bool = False):
    """--topk 0: (disp, margin, tie_mask or None)."""
    L = np.ascontiguousarray(L, np.uint8)
    R = np.ascontiguousarray(R, np.uint8)
    H, W = L.shape
    disp = np.empty((H, W), np.float32)
    margin = np.empty((H, W), np.float32)
    ties = np.zeros((H, W), np.uint8) if want_ties else None
    rc = lib().orc_dense_run_full(_p(L), _p(R), W, H, C.byref(c), _p(disp), _p(margin),
                                  _p(ties) if ties is not None else None)
    if rc:
        raise ValueError(f"orc_dense_run_full rc={rc}")
    return disp, margin, ties


def volume_f32(L, R, c: OrcCfg) -> np.ndarray:
    L = np.ascontiguousarray(L, np.uint8)
    R = np.ascontiguousarray(R, np.uint8)
    H, W = L.shape
    D = c.dmax - c.dmin + 1
    vol = np.empty((H, W, D), np.float32)
    rc = lib().orc_volume_f32(_p(L), _p(R), W, H, C.byref(c), _p(vol))
    if rc:
        raise ValueError(f"orc_volume_f32 rc={rc}")
    return vol


def confidence(s1, s2, lam, lrc=0.0, tol=1.0) -> float:
    return float(lib().orc_confidence(s1, s2, lam, lrc, tol))


def confidence_u8(s1, s2, lam, lrc=0.0, tol=1.0) -> int:
    return int(lib().orc_confidence_u8(s1, s2, lam, lrc, tol))


def dedd_planes(disp, cand6, lam: float = -0.1):
    """Q4 disparity, u8 confidence and the float P*255 it rounds, from a
    frame's disparity and --out-conf planes (s1 = plane 0, s2 = plane 1, the
    solver's reverse-match distance = plane 5), as the reference's stream
    output computes them."""
    disp = np.ascontiguousarray(disp, np.float32)
    c6 = np.ascontiguousarray(cand6, np.float32)
    n = disp.size
    q4 = np.empty(disp.shape, np.int16)
    conf = np.empty(disp.shape, np.uint8)
    p255 = np.empty(disp.shape, np.float32)
    s1, s2, lrc = (np.ascontiguousarray(c6[i]) for i in (0, 1, 5))
    lib().orc_dedd_planes(_p(disp), _p(s1), _p(s2), _p(lrc), n, lam, _p(q4), _p(conf),
                          _p(p255))
    return q4, conf, p255
