"""ctypes loader for libdoubleeye_hip.so (the C ABI in include/doubleeye_hip.h).

The library is the product path: if it is missing or cannot be loaded this
module raises; there is no fallback implementation.
"""
from __future__ import annotations

import ctypes as C
import os
import threading

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libdoubleeye_hip.so")

DE_OK = 0
DE_ERR_ARG = -1
DE_ERR_UNSUPPORTED = -2
DE_ERR_HIP = -3
DE_ERR_NOMEM = -4
DE_ERR_BUSY = -5


class DeError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"doubleeye_hip error {code}: {msg}")
        self.code = code


class DeCfgC(C.Structure):
    """de_cfg (include/doubleeye_hip.h)."""
    _fields_ = [("dmin", C.c_int), ("dmax", C.c_int), ("iters", C.c_int),
                ("census_hw", C.c_int), ("census_hh", C.c_int),
                ("sigma_s", C.c_float), ("sigma_r", C.c_float),
                ("ad", C.c_float), ("ad_trunc", C.c_int),
                ("lambda_", C.c_float), ("gamma", C.c_float),
                ("damping", C.c_float), ("min_margin", C.c_float),
                ("min_ratio", C.c_float),
                ("subpixel", C.c_int), ("fit_eq", C.c_int), ("lrc", C.c_int)]


class DePairC(C.Structure):
    """de_pair (include/doubleeye_hip.h): device pointers of one batch pair."""
    _fields_ = [("L", C.c_void_p), ("R", C.c_void_p), ("W", C.c_int), ("H", C.c_int),
                ("disp", C.c_void_p), ("margin", C.c_void_p), ("cand6", C.c_void_p),
                ("conf", C.c_void_p), ("q4", C.c_void_p)]


class DeMatchCfgC(C.Structure):
    """de_match_cfg (include/doubleeye_hip.h)."""
    _fields_ = [("max_dy", C.c_float), ("min_disparity", C.c_float),
                ("max_disparity", C.c_float), ("max_candidates", C.c_int),
                ("sigma_y", C.c_float), ("w_y", C.c_float), ("w_prior", C.c_float),
                ("w_smooth", C.c_float), ("iterations", C.c_int), ("damping", C.c_float),
                ("lambda_", C.c_float), ("gamma", C.c_float), ("converge_eps", C.c_float)]


class DePriorC(C.Structure):
    """de_disparity_prior (include/doubleeye_hip.h)."""
    _fields_ = [("cell", C.c_int), ("cols", C.c_int), ("rows", C.c_int),
                ("disparity", C.c_void_p), ("sigma", C.c_void_p), ("valid", C.c_void_p)]


class DeSmoothC(C.Structure):
    """de_smooth_prior (include/doubleeye_hip.h)."""
    _fields_ = [("n", C.c_int), ("disparity", C.c_void_p), ("sigma", C.c_void_p),
                ("valid", C.c_void_p)]


class DeMatchStatsC(C.Structure):
    """de_match_stats (include/doubleeye_hip.h)."""
    _fields_ = [("left", C.c_int), ("right", C.c_int), ("candidates", C.c_int),
                ("iterations_run", C.c_int), ("converged", C.c_int),
                ("final_max_delta", C.c_float), ("oscillating", C.c_int)]


class DeSparseFrameC(C.Structure):
    """de_sparse_frame (include/doubleeye_hip.h): device pointers of one frame."""
    _fields_ = [("left", C.c_void_p), ("desc_left", C.c_void_p), ("n_left", C.c_int),
                ("right", C.c_void_p), ("desc_right", C.c_void_p), ("n_right", C.c_int),
                ("out", C.c_void_p), ("n_out", C.c_void_p), ("stats", C.c_void_p)]


class DeKnnCfgC(C.Structure):
    """de_knn_cfg (include/doubleeye_hip.h)."""
    _fields_ = [("census_hw", C.c_int), ("census_hh", C.c_int), ("dmax", C.c_int),
                ("n_cand", C.c_int), ("k_nn", C.c_int), ("iters", C.c_int),
                ("w_data", C.c_float), ("lambda_", C.c_float), ("trunc", C.c_int)]


class DeKnnFrameC(C.Structure):
    """de_knn_frame (include/doubleeye_hip.h): device pointers of one frame."""
    _fields_ = [("L", C.c_void_p), ("R", C.c_void_p), ("W", C.c_int), ("H", C.c_int),
                ("kp", C.c_void_p), ("n", C.c_int), ("out", C.c_void_p)]


class DeServerCfgC(C.Structure):
    """de_server_cfg (include/doubleeye_hip.h)."""
    _fields_ = [("mem_budget", C.c_size_t), ("hip_streams", C.c_int), ("slots", C.c_int),
                ("max_batch", C.c_int)]


class DeServerResultC(C.Structure):
    """de_server_result (include/doubleeye_hip.h)."""
    _fields_ = [("stream_id", C.c_int), ("slot", C.c_int), ("kind", C.c_int), ("status", C.c_int),
                ("tag", C.c_uint64), ("t_submit", C.c_double), ("t_start", C.c_double),
                ("t_done", C.c_double), ("batch_size", C.c_int), ("W", C.c_int), ("H", C.c_int),
                ("n", C.c_int), ("disp", C.c_void_p), ("q4", C.c_void_p), ("conf", C.c_void_p),
                ("knn", C.c_void_p)]


class DeDetectorCfgC(C.Structure):
    """de_detector_cfg (include/doubleeye_hip.h)."""
    _fields_ = [("grad_window", C.c_int), ("cell", C.c_int), ("per_cell", C.c_int),
                ("min_response", C.c_float), ("min_local_std", C.c_float),
                ("border", C.c_int), ("subpixel", C.c_int), ("fast_threshold", C.c_int),
                ("nms_radius", C.c_int)]


class DeDetectProfileC(C.Structure):
    """de_detect_profile (include/doubleeye_hip.h)."""
    _fields_ = [("fast_ms", C.c_double), ("nms_ms", C.c_double), ("refine_ms", C.c_double),
                ("candidates", C.c_int), ("suppressed", C.c_int)]


class DeDetectFrameC(C.Structure):
    """de_detect_frame (include/doubleeye_hip.h): device pointers of one image."""
    _fields_ = [("img", C.c_void_p), ("W", C.c_int), ("H", C.c_int), ("out", C.c_void_p),
                ("cap", C.c_int), ("n_out", C.c_void_p)]


# (name, restype, argtypes) for every exported entry point
P = C.c_void_p
I = C.c_int
F = C.c_float
SIGNATURES = {
    "de_default_cfg": (None, [C.POINTER(DeCfgC)]),
    "de_ctx_create": (I, [I, C.POINTER(P)]),
    "de_ctx_destroy": (None, [P]),
    "de_ctx_sync": (I, [P]),
    "de_ctx_stream": (P, [P]),
    "de_last_error": (C.c_char_p, []),
    "de_alloc": (I, [P, C.c_size_t, C.POINTER(P)]),
    "de_free": (I, [P, P]),
    "de_upload": (I, [P, P, P, C.c_size_t]),
    "de_download": (I, [P, P, P, C.c_size_t]),
    "de_copy_dev": (I, [P, P, P, C.c_size_t]),
    "de_fill_dev": (I, [P, P, I, C.c_size_t]),
    "de_census": (I, [P, P, I, I, I, I, P]),
    "de_census_dev": (I, [P, P, I, I, I, I, P, P]),
    "de_dense_run": (I, [P, P, P, I, I, C.POINTER(DeCfgC), P, P, P]),
    "de_host_alloc": (I, [C.c_size_t, C.POINTER(C.c_void_p)]),
    "de_host_free": (I, [P]),
    "de_dense_run_dev": (I, [P, P, P, I, I, C.POINTER(DeCfgC), P, P, P, P]),
    "de_dense_run_dev_noalloc": (I, [P, P, P, I, I, C.POINTER(DeCfgC), P, P, P, P]),
    "de_dense_noalloc_limit": (C.c_size_t, []),
    "de_dense_volume": (I, [P, P, P, I, I, C.POINTER(DeCfgC), I, I, P]),
    "de_dense_candidates": (I, [P, P, P, I, I, C.POINTER(DeCfgC), P]),
    "de_ctx_set_profiling": (I, [P, I]),
    "de_ctx_set_variants": (I, [P, I, I]),
    "de_ctx_stage_ms": (I, [P, C.POINTER(C.c_double), C.POINTER(C.c_int)]),
    "de_dense_batch_dev": (I, [P, C.POINTER(DePairC), C.POINTER(DeCfgC), I, P]),
    "de_plan_create": (I, [P, C.POINTER(DePairC), C.POINTER(DeCfgC), I, C.POINTER(P)]),
    "de_plan_launch": (I, [P, P]),
    "de_plan_destroy": (None, [P]),
    "de_plan_parts": (I, [P]),
    "de_autotune": (I, [P, I, I, C.POINTER(DeCfgC), I, C.POINTER(I), C.POINTER(I)]),
    "de_match_default_cfg": (None, [C.POINTER(DeMatchCfgC)]),
    "de_match_masda": (I, [P, P, P, I, P, P, I, C.POINTER(DeMatchCfgC), P, P, P,
                           C.POINTER(I), P]),
    "de_match_mutual_nn": (I, [P, P, P, I, P, P, I, C.POINTER(DeMatchCfgC), P, P, F, P,
                               C.POINTER(I), P]),
    "de_match_candidates": (I, [P, P, P, I, P, P, I, C.POINTER(DeMatchCfgC), P, P, P, P, P,
                                C.POINTER(I)]),
    "de_match_masda_batch_dev": (I, [P, C.POINTER(DeSparseFrameC), C.POINTER(DeMatchCfgC),
                                     I, P]),
    "de_knn_default_cfg": (None, [C.POINTER(DeKnnCfgC)]),
    "de_knn_bp": (I, [P, P, P, I, I, P, I, C.POINTER(DeKnnCfgC), P]),
    "de_knn_candidates": (I, [P, P, P, I, I, P, I, C.POINTER(DeKnnCfgC), P, P, P]),
    "de_knn_graph": (I, [P, P, P, I, I, P, I, C.POINTER(DeKnnCfgC), P, P, I, C.POINTER(I)]),
    "de_knn_bp_graph": (I, [P, I, P, P, P, P, P, C.POINTER(DeKnnCfgC), P, P]),
    "de_knn_bp_batch_dev": (I, [P, C.POINTER(DeKnnFrameC), C.POINTER(DeKnnCfgC), I, P]),
    "de_dense_volume_f32": (I, [P, P, P, I, I, C.POINTER(DeCfgC), P]),
    "de_dense_run_full": (I, [P, P, P, I, I, C.POINTER(DeCfgC), P, P]),
    "de_dense_run_full_dev": (I, [P, P, P, I, I, C.POINTER(DeCfgC), P, P, P]),
    "de_dense_problem": (I, [P, P, P, I, I, I, I, I, I, P, P, P, C.c_int64,
                             C.POINTER(C.c_int64)]),
    "de_masda_edges": (I, [P, P, P, P, C.c_int64, C.c_int64, C.c_int64, C.c_double, C.c_double,
                           I, C.c_double, C.c_double, P, P, C.POINTER(I)]),
    "de_masda_edges_f32": (I, [P, P, P, P, C.c_int64, C.c_int64, C.c_int64, C.c_double,
                               C.c_double, I, C.c_double, C.c_double, P, P, C.POINTER(I)]),
    "de_masda_matrix": (I, [P, P, C.c_int64, C.c_int64, C.c_double, C.c_double, I, C.c_double,
                            C.c_double, P, P, C.POINTER(I)]),
    "de_masda_decide": (I, [P, P, P, P, P, C.c_int64, C.c_int64, C.c_int64, C.c_double, P]),
    "de_seg_max_excluding": (I, [P, P, P, C.c_int64, C.c_int64, P]),
    "de_census_points": (I, [P, P, I, I, P, C.c_int64, I, I, P]),
    "de_hamming": (I, [P, P, P, C.c_int64, P]),
    "de_build_problem": (I, [P, P, P, C.c_int64, P, P, C.c_int64, I, C.c_double, C.c_double,
                             C.c_double, C.c_double, C.c_double, P]),
    "de_dense_masda": (I, [P, P, P, I, I, I, I, I, I, C.c_double, C.c_double, I, C.c_double,
                           C.c_double, P, P, C.POINTER(I)]),
    "de_stream_packet_bytes": (C.c_size_t, [I, I, I]),
    "de_stream_create": (I, [P, I, I, C.POINTER(DeCfgC), I, C.POINTER(P)]),
    "de_stream_destroy": (None, [P]),
    "de_stream_feed": (I, [P, P, C.c_size_t, P, C.c_size_t, C.POINTER(C.c_size_t)]),
    "de_stream_stats": (I, [P, C.POINTER(C.c_long), C.POINTER(C.c_long)]),
    "de_stream_run_fd": (I, [P, I, I, I, I, C.POINTER(DeCfgC), I, C.POINTER(C.c_long),
                             C.POINTER(C.c_long)]),
    "de_server_default_cfg": (None, [C.POINTER(DeServerCfgC)]),
    "de_server_create": (I, [P, C.POINTER(DeServerCfgC), C.POINTER(P)]),
    "de_server_destroy": (None, [P]),
    "de_server_open_dense": (I, [P, I, I, C.POINTER(DeCfgC), C.POINTER(I)]),
    "de_server_open_knn": (I, [P, I, I, I, C.POINTER(DeKnnCfgC), C.POINTER(I)]),
    "de_server_close": (I, [P, I]),
    "de_server_memory": (I, [P, C.POINTER(C.c_size_t), C.POINTER(C.c_size_t)]),
    "de_server_submit": (I, [P, I, P, P, P, I, C.c_uint64, C.POINTER(I)]),
    "de_server_poll": (I, [P, C.POINTER(DeServerResultC), I, I, C.POINTER(I)]),
    "de_server_release": (I, [P, I, I]),
    "de_dense_run_prior": (I, [P, P, P, I, I, C.POINTER(DeCfgC), P, C.c_size_t, I, P, P, P]),
    "de_dense_run_c2f": (I, [P, P, P, I, I, C.POINTER(DeCfgC), I, P, P, P, P, C.POINTER(I)]),
    "de_dense_run_c2f_dev": (I, [P, P, P, I, I, C.POINTER(DeCfgC), I, P, P, P, C.POINTER(I), P]),
    "de_detector_default_cfg": (None, [C.POINTER(DeDetectorCfgC)]),
    "de_detect_max_keypoints": (C.c_size_t, [I, I, C.POINTER(DeDetectorCfgC)]),
    "de_detect_keypoints_fast": (I, [P, P, I, I, C.POINTER(DeDetectorCfgC), P, I, C.POINTER(I),
                                     P]),
    "de_detect_keypoints_fast_batch_dev": (I, [P, C.POINTER(DeDetectFrameC),
                                               C.POINTER(DeDetectorCfgC), I, P]),
    "de_fast_score_map": (I, [P, P, I, I, I, I, P, P]),
    "de_census_at_keypoints": (I, [P, P, I, I, I, I, P, I, P]),
    "de_refine_disparity": (I, [P, P, P, I, I, P, I, P, I, P, I, I]),
    "de_sparse_pipeline": (I, [P, P, P, I, I, C.POINTER(DeDetectorCfgC), P, I, I,
                               C.POINTER(DeMatchCfgC), F, I, P, I, C.POINTER(I), P, I,
                               C.POINTER(I), P, P, I, C.POINTER(I), P]),
}

_lib = None
_lock = threading.Lock()


def lib() -> C.CDLL:
    """Load the HIP library (raises if it has not been built)."""
    global _lib
    with _lock:
        if _lib is None:
            if not os.path.exists(LIB_PATH):
                raise ImportError(
                    f"{LIB_PATH} is missing: build it with `python -m doubleeye_amd.build` "
                    "(there is no CPU fallback)")
            L = C.CDLL(LIB_PATH)
            for name, (res, args) in SIGNATURES.items():
                fn = getattr(L, name, None)
                if fn is None:
                    continue
                fn.restype = res
                fn.argtypes = args
            _lib = L
    return _lib


def check(rc: int) -> None:
    if rc != DE_OK:
        msg = lib().de_last_error()
        raise DeError(rc, msg.decode() if msg else "")
