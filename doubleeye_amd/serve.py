"""Serving loop over the C ABI: many camera streams of mixed geometry, dense
or keypoint-graph, on one GPU with continuous batching (include/doubleeye_hip.h,
"Serving loop"). The scheduler is native (a thread inside the library); this
module opens streams, submits frames and collects results.

The reference's serving analogue is one fixed-geometry stream with one frame in
flight (`de_dense_cuda --stream`, core/tools/de_dense_cuda.cu:670-770).
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from ._lib import (DE_ERR_BUSY, DeServerCfgC, DeServerResultC, DeError, check, lib)
from .dense import Cfg
from .knn import OUT as KNN_OUT
from .knn import KnnConfig

__all__ = ["Server", "Result", "DENSE", "KNN"]

DENSE, KNN = 0, 1


@dataclass
class Result:
    stream_id: int
    slot: int
    kind: int
    status: int
    tag: int
    t_submit: float
    t_start: float
    t_done: float
    batch_size: int
    W: int
    H: int
    n: int
    _raw: DeServerResultC = None

    @property
    def latency(self) -> float:
        return self.t_done - self.t_submit

    def fetch(self):
        """Copy the result out of the slot: dense (disp, q4, conf), knn OUT array."""
        r = self._raw
        if self.kind == DENSE:
            n = self.W * self.H
            disp = np.ctypeslib.as_array(C.cast(r.disp, C.POINTER(C.c_float)), (n,)).copy()
            q4 = np.ctypeslib.as_array(C.cast(r.q4, C.POINTER(C.c_int16)), (n,)).copy()
            conf = np.ctypeslib.as_array(C.cast(r.conf, C.POINTER(C.c_uint8)), (n,)).copy()
            return (disp.reshape(self.H, self.W), q4.reshape(self.H, self.W),
                    conf.reshape(self.H, self.W))
        if self.n == 0:
            return np.zeros(0, KNN_OUT)
        buf = (C.c_uint8 * (self.n * KNN_OUT.itemsize)).from_address(r.knn)
        return np.frombuffer(bytes(buf), KNN_OUT).copy()


class Server:
    def __init__(self, ctx, mem_budget: int = 8 << 30, hip_streams: int = 4, slots: int = 4,
                 max_batch: int = 0):
        self.ctx = ctx
        c = DeServerCfgC(mem_budget, hip_streams, slots, max_batch)
        h = C.c_void_p()
        check(lib().de_server_create(ctx.h, C.byref(c), C.byref(h)))
        self.h = h.value
        self._res = (DeServerResultC * 256)()

    def close(self):
        if self.h:
            lib().de_server_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def open_dense(self, W: int, H: int, cfg: Cfg) -> int:
        sid = C.c_int(-1)
        check(lib().de_server_open_dense(self.h, W, H, C.byref(cfg.to_c()), C.byref(sid)))
        return sid.value

    def open_knn(self, W: int, H: int, max_kp: int, cfg: KnnConfig | None = None) -> int:
        sid = C.c_int(-1)
        check(lib().de_server_open_knn(self.h, W, H, max_kp, C.byref((cfg or KnnConfig()).to_c()),
                                       C.byref(sid)))
        return sid.value

    def close_stream(self, sid: int):
        check(lib().de_server_close(self.h, sid))

    def memory(self):
        u, b = C.c_size_t(), C.c_size_t()
        check(lib().de_server_memory(self.h, C.byref(u), C.byref(b)))
        return u.value, b.value

    def submit(self, sid: int, L: np.ndarray, R: np.ndarray, kp: np.ndarray | None = None,
               tag: int = 0):
        """Returns the slot, or None when the stream's slots are all taken."""
        L = np.ascontiguousarray(L, np.uint8)
        R = np.ascontiguousarray(R, np.uint8)
        k = None if kp is None else np.ascontiguousarray(kp, np.int32).reshape(-1, 2)
        slot = C.c_int(-1)
        rc = lib().de_server_submit(self.h, sid, L.ctypes.data_as(C.c_void_p),
                                    R.ctypes.data_as(C.c_void_p),
                                    k.ctypes.data_as(C.c_void_p) if k is not None and len(k) else None,
                                    0 if k is None else len(k), tag, C.byref(slot))
        if rc == DE_ERR_BUSY:
            return None
        check(rc)
        return slot.value

    def poll(self, timeout_ms: int = 10, max_results: int = 256):
        n = C.c_int(0)
        check(lib().de_server_poll(self.h, self._res, min(max_results, 256), timeout_ms,
                                   C.byref(n)))
        out = []
        for i in range(n.value):
            r = self._res[i]
            raw = DeServerResultC()
            C.pointer(raw)[0] = r
            out.append(Result(r.stream_id, r.slot, r.kind, r.status, r.tag, r.t_submit, r.t_start,
                              r.t_done, r.batch_size, r.W, r.H, r.n, raw))
        return out

    def release(self, res: Result):
        check(lib().de_server_release(self.h, res.stream_id, res.slot))


def percentile(xs, p):
    xs = sorted(xs)
    if not xs:
        return float("nan")
    k = min(len(xs) - 1, max(0, int(round(p / 100.0 * (len(xs) - 1)))))
    return xs[k]
