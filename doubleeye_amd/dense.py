"""Dense MASDA on the MI355X through the C ABI.

Mirrors the reference's dense entry points:

* ``Cfg`` -- the default-path fields of ``doubleeye::Cfg``
  (core/include/doubleeye/dense_solve.hpp:19-97), same names and defaults.
* ``run_dense(L, R, cfg)`` -- ``run_dense`` (core/tools/de_dense.cpp:738) on its
  default path, i.e. what ``de_dense --threads 1`` writes with ``--out`` and
  ``--out-conf``.
* ``census(img, hw, hh)`` -- ``census_transform`` (core/src/preproc.cpp:110).

Errors are raised as ``DeError`` (the reference exits 1/2 with a message).
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, fields

import numpy as np

from ._lib import DeCfgC, DePairC, check, lib

__all__ = ["Cfg", "Context", "DeviceArray", "Pair", "Plan", "run_dense", "census",
           "default_context", "host_empty"]


def host_empty(shape, dtype=np.uint8) -> np.ndarray:
    """An uninitialised numpy array in pinned host memory (de_host_alloc).

    The host-pointer entry points move such arrays by DMA alone; ordinary
    arrays go through the context's pinned staging. The memory is returned
    with de_host_free when the last view of the array is gone."""
    import weakref

    dt = np.dtype(dtype)
    n = max(1, int(np.prod(shape, dtype=np.int64)) * dt.itemsize)
    p = C.c_void_p()
    check(lib().de_host_alloc(n, C.byref(p)))
    buf = (C.c_uint8 * n).from_address(p.value)
    weakref.finalize(buf, lib().de_host_free, p.value)
    return np.frombuffer(buf, np.uint8, int(np.prod(shape, dtype=np.int64)) * dt.itemsize) \
        .view(dt).reshape(shape)


@dataclass
class Cfg:
    dmin: int = 1
    dmax: int = 60
    iters: int = 0
    census_hw: int = 3
    census_hh: int = 3
    sigma_s: float = 8.0
    sigma_r: float = 0.2
    ad: float = 0.15
    ad_trunc: int = 10
    lam: float = -0.1          # lambda (a Python keyword)
    gamma: float = -0.1
    damping: float = 0.4
    min_margin: float = 0.0
    min_ratio: float = 0.0
    subpixel: bool = True
    fit_eq: bool = True
    lrc: bool = False

    @property
    def D(self) -> int:
        return self.dmax - self.dmin + 1

    def to_c(self) -> DeCfgC:
        c = DeCfgC()
        for f in fields(self):
            name = "lambda_" if f.name == "lam" else f.name
            v = getattr(self, f.name)
            setattr(c, name, int(v) if isinstance(v, bool) else v)
        return c


class DeviceArray:
    """Device memory owned through the library's HIP runtime (de_alloc), with a
    numpy dtype/shape for uploads and downloads."""

    def __init__(self, ctx: "Context", shape, dtype):
        self.ctx = ctx
        self.shape = tuple(shape)
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape)) * self.dtype.itemsize
        h = C.c_void_p()
        check(lib().de_alloc(ctx.h, self.nbytes, C.byref(h)))
        self.ptr = int(h.value or 0)

    @classmethod
    def from_numpy(cls, ctx: "Context", a: np.ndarray) -> "DeviceArray":
        a = np.ascontiguousarray(a)
        d = cls(ctx, a.shape, a.dtype)
        d.upload(a)
        return d

    def upload(self, a: np.ndarray):
        a = np.ascontiguousarray(a, self.dtype)
        if a.nbytes != self.nbytes:
            raise ValueError(f"upload of {a.nbytes} bytes into {self.nbytes}")
        check(lib().de_upload(self.ctx.h, C.c_void_p(self.ptr), a.ctypes.data_as(C.c_void_p),
                              self.nbytes))

    def numpy(self) -> np.ndarray:
        out = np.empty(self.shape, self.dtype)
        check(lib().de_download(self.ctx.h, out.ctypes.data_as(C.c_void_p), C.c_void_p(self.ptr),
                                self.nbytes))
        return out

    def free(self):
        if self.ptr and self.ctx.h:
            lib().de_free(self.ctx.h, C.c_void_p(self.ptr))
        self.ptr = 0

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


@dataclass
class Pair:
    """One pair of a batch: device addresses (ints) of the inputs and outputs."""
    L: int
    R: int
    W: int
    H: int
    disp: int
    margin: int
    cand6: int = 0
    conf: int = 0
    q4: int = 0

    def to_c(self) -> DePairC:
        return DePairC(self.L, self.R, self.W, self.H, self.disp, self.margin,
                       self.cand6 or None, self.conf or None, self.q4 or None)


def _pairs_cfgs(pairs, cfgs):
    n = len(pairs)
    if len(cfgs) != n:
        raise ValueError(f"{n} pairs but {len(cfgs)} configurations")
    P = (DePairC * n)(*[p.to_c() for p in pairs])
    Cs = (DeCfgC * n)(*[c.to_c() for c in cfgs])
    return P, Cs, n


class Plan:
    """A captured batch (de_plan): resident scratch + a hipGraph of the stages."""

    def __init__(self, ctx: "Context", pairs, cfgs):
        P, Cs, n = _pairs_cfgs(pairs, cfgs)
        h = C.c_void_p()
        check(lib().de_plan_create(ctx.h, P, Cs, n, C.byref(h)))
        self.h = h
        self.n = n

    def launch(self, stream: int = 0):
        check(lib().de_plan_launch(self.h, C.c_void_p(stream) if stream else None))

    @property
    def parts(self) -> int:
        """Sub-batches the plan pipelines (1: one graph of every stage)."""
        n = lib().de_plan_parts(self.h)
        if n < 0:
            check(n)
        return n

    def close(self):
        if self.h:
            lib().de_plan_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _p(a: np.ndarray):
    return a.ctypes.data_as(C.c_void_p)


class Context:
    """A device context: stream + resident scratch (de_ctx)."""

    def __init__(self, device: int = 0):
        h = C.c_void_p()
        check(lib().de_ctx_create(device, C.byref(h)))
        self.h = h
        self.device = device

    def close(self):
        if self.h:
            lib().de_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def stream(self) -> int:
        return int(lib().de_ctx_stream(self.h) or 0)

    def sync(self):
        check(lib().de_ctx_sync(self.h))

    # --- host-pointer API ---
    def census(self, img: np.ndarray, hw: int = 3, hh: int = 3) -> np.ndarray:
        img = np.ascontiguousarray(img, np.uint8)
        out = np.empty(img.shape, np.uint64)
        check(lib().de_census(self.h, _p(img), img.shape[1], img.shape[0], hw, hh, _p(out)))
        return out

    def run_dense(self, L: np.ndarray, R: np.ndarray, cfg: Cfg | None = None,
                  want_cand: bool = False, out=None):
        """(disp, margin, cand6 or None). `out`: (disp, margin[, cand6]) arrays to
        write into (float32, C-contiguous, H x W and 6 x H x W), e.g. from
        host_empty, instead of new ones."""
        cfg = cfg or Cfg()
        L = np.ascontiguousarray(L, np.uint8)
        R = np.ascontiguousarray(R, np.uint8)
        if L.shape != R.shape or L.ndim != 2:
            raise ValueError(f"left {L.shape} and right {R.shape} must be equal 2-D images")
        H, W = L.shape
        if out is not None:
            disp, margin = out[0], out[1]
            cand = out[2] if want_cand else None
            for a, shp in ((disp, (H, W)), (margin, (H, W)), (cand, (6, H, W))):
                if a is not None and (a.dtype != np.float32 or a.shape != shp
                                      or not a.flags.c_contiguous):
                    raise ValueError(f"out array {a.dtype} {a.shape}: expected float32 {shp}, "
                                     "C-contiguous")
        else:
            disp = np.empty((H, W), np.float32)
            margin = np.empty((H, W), np.float32)
            cand = np.empty((6, H, W), np.float32) if want_cand else None
        c = cfg.to_c()
        check(lib().de_dense_run(self.h, _p(L), _p(R), W, H, C.byref(c), _p(disp), _p(margin),
                                 _p(cand) if cand is not None else None))
        return disp, margin, cand

    def run_dense_prior(self, L, R, prior, cfg: Cfg | None = None, band: int = 2,
                        want_cand: bool = False):
        """`de_dense --prior F --band N`: (disp, margin, cand6 or None). `prior` is
        H x W, or H/2 x W/2 (upsampled nearest, disparity doubled)."""
        cfg = cfg or Cfg()
        L = np.ascontiguousarray(L, np.uint8)
        R = np.ascontiguousarray(R, np.uint8)
        H, W = L.shape
        pr = np.ascontiguousarray(prior, np.float32)
        disp = np.empty((H, W), np.float32)
        margin = np.empty((H, W), np.float32)
        cand = np.empty((6, H, W), np.float32) if want_cand else None
        check(lib().de_dense_run_prior(self.h, _p(L), _p(R), W, H, C.byref(cfg.to_c()), _p(pr),
                                       pr.size, int(band), _p(disp), _p(margin),
                                       _p(cand) if cand is not None else None))
        return disp, margin, cand

    def run_dense_c2f(self, L, R, cfg: Cfg | None = None, band: int = 2,
                      want_cand: bool = False):
        """`de_dense --c2f --band N`: (disp, margin, cand6 or None, the filled
        coarse map, whether the masked sweep ran)."""
        cfg = cfg or Cfg()
        L = np.ascontiguousarray(L, np.uint8)
        R = np.ascontiguousarray(R, np.uint8)
        H, W = L.shape
        disp = np.empty((H, W), np.float32)
        margin = np.empty((H, W), np.float32)
        cand = np.empty((6, H, W), np.float32) if want_cand else None
        coarse = np.empty((H // 2, W // 2), np.float32)
        used = C.c_int(0)
        check(lib().de_dense_run_c2f(self.h, _p(L), _p(R), W, H, C.byref(cfg.to_c()), int(band),
                                     _p(disp), _p(margin), _p(cand) if cand is not None else None,
                                     _p(coarse), C.byref(used)))
        return disp, margin, cand, coarse, bool(used.value)

    def run_dense_c2f_dev(self, dL: int, dR: int, W: int, H: int, cfg: Cfg, d_disp: int,
                          d_margin: int, d_cand6: int = 0, band: int = 2,
                          stream: int = 0) -> bool:
        """de_dense_run_c2f_dev: --c2f on device addresses, queued on the
        stream after one wait for the coarse level; True when the masked sweep
        ran."""
        used = C.c_int(0)
        check(lib().de_dense_run_c2f_dev(self.h, C.c_void_p(dL), C.c_void_p(dR), W, H,
                                         C.byref(cfg.to_c()), int(band), C.c_void_p(d_disp),
                                         C.c_void_p(d_margin),
                                         C.c_void_p(d_cand6) if d_cand6 else None,
                                         C.byref(used), C.c_void_p(stream) if stream else None))
        return bool(used.value)

    def volume(self, L, R, cfg: Cfg, k0: int, k1: int) -> np.ndarray:
        L = np.ascontiguousarray(L, np.uint8)
        R = np.ascontiguousarray(R, np.uint8)
        H, W = L.shape
        out = np.empty((k1 - k0, H, W), np.int16)
        c = cfg.to_c()
        check(lib().de_dense_volume(self.h, _p(L), _p(R), W, H, C.byref(c), k0, k1, _p(out)))
        return out

    def volume_f32(self, L, R, cfg: Cfg) -> np.ndarray:
        """The FP32 aggregated volume [y][x][k] (`de_dense --dump-vol`)."""
        L = np.ascontiguousarray(L, np.uint8)
        R = np.ascontiguousarray(R, np.uint8)
        H, W = L.shape
        out = np.empty((H, W, cfg.D), np.float32)
        check(lib().de_dense_volume_f32(self.h, _p(L), _p(R), W, H, C.byref(cfg.to_c()), _p(out)))
        return out

    def run_dense_full(self, L, R, cfg: Cfg):
        """`de_dense --topk 0`: FP32 volume + full-volume row solver ->
        (disp, margin)."""
        L = np.ascontiguousarray(L, np.uint8)
        R = np.ascontiguousarray(R, np.uint8)
        H, W = L.shape
        disp = np.empty((H, W), np.float32)
        margin = np.empty((H, W), np.float32)
        check(lib().de_dense_run_full(self.h, _p(L), _p(R), W, H, C.byref(cfg.to_c()), _p(disp),
                                      _p(margin)))
        return disp, margin

    def run_dense_full_dev(self, dL: int, dR: int, W: int, H: int, cfg: Cfg, d_disp: int,
                           d_margin: int, stream: int = 0):
        """de_dense_run_full_dev: `--topk 0` on device addresses, queued on the
        stream (null: the context's)."""
        check(lib().de_dense_run_full_dev(self.h, C.c_void_p(dL), C.c_void_p(dR), W, H,
                                          C.byref(cfg.to_c()), C.c_void_p(d_disp),
                                          C.c_void_p(d_margin),
                                          C.c_void_p(stream) if stream else None))

    def candidates(self, L, R, cfg: Cfg) -> np.ndarray:
        L = np.ascontiguousarray(L, np.uint8)
        R = np.ascontiguousarray(R, np.uint8)
        H, W = L.shape
        out = np.empty((H, W, 4), np.uint32)
        c = cfg.to_c()
        check(lib().de_dense_candidates(self.h, _p(L), _p(R), W, H, C.byref(c), _p(out)))
        return out

    # --- device-pointer API (ints are device addresses, e.g. tensor.data_ptr()) ---
    def run_dense_dev(self, dL: int, dR: int, W: int, H: int, cfg: Cfg, d_disp: int,
                      d_margin: int, d_cand6: int = 0, stream: int = 0):
        c = cfg.to_c()
        check(lib().de_dense_run_dev(self.h, C.c_void_p(dL), C.c_void_p(dR), W, H, C.byref(c),
                                     C.c_void_p(d_disp), C.c_void_p(d_margin),
                                     C.c_void_p(d_cand6) if d_cand6 else None,
                                     C.c_void_p(stream) if stream else None))

    def run_dense_dev_noalloc(self, dL: int, dR: int, W: int, H: int, cfg: Cfg, d_disp: int,
                              d_margin: int, d_cand6: int = 0, stream: int = 0):
        """run_dense_dev with no device allocation (the library's static scratch;
        calls that overlap on one device race)."""
        c = cfg.to_c()
        check(lib().de_dense_run_dev_noalloc(self.h, C.c_void_p(dL), C.c_void_p(dR), W, H,
                                             C.byref(c), C.c_void_p(d_disp),
                                             C.c_void_p(d_margin),
                                             C.c_void_p(d_cand6) if d_cand6 else None,
                                             C.c_void_p(stream) if stream else None))

    def census_dev(self, d_img: int, W: int, H: int, hw: int, hh: int, d_out: int,
                   stream: int = 0):
        check(lib().de_census_dev(self.h, C.c_void_p(d_img), W, H, hw, hh, C.c_void_p(d_out),
                                  C.c_void_p(stream) if stream else None))

    def empty(self, shape, dtype) -> DeviceArray:
        return DeviceArray(self, shape, dtype)

    def to_device(self, a: np.ndarray) -> DeviceArray:
        return DeviceArray.from_numpy(self, a)

    def copy_dev(self, dst: DeviceArray, src: DeviceArray):
        check(lib().de_copy_dev(self.h, C.c_void_p(dst.ptr), C.c_void_p(src.ptr),
                                min(dst.nbytes, src.nbytes)))

    def fill_dev(self, dst: DeviceArray, byte: int):
        check(lib().de_fill_dev(self.h, C.c_void_p(dst.ptr), int(byte), dst.nbytes))

    # --- batches (device pointers) ---
    def dense_batch_dev(self, pairs, cfgs, stream: int = 0):
        P, Cs, n = _pairs_cfgs(pairs, cfgs)
        check(lib().de_dense_batch_dev(self.h, P, Cs, n, C.c_void_p(stream) if stream else None))

    def plan(self, pairs, cfgs) -> Plan:
        return Plan(self, pairs, cfgs)

    def autotune(self, W: int, H: int, cfg: Cfg, reps: int = 5):
        rv, cv = C.c_int(), C.c_int()
        c = cfg.to_c()
        check(lib().de_autotune(self.h, W, H, C.byref(c), reps, C.byref(rv), C.byref(cv)))
        return rv.value, cv.value

    def set_variants(self, row: int = -1, col: int = -1):
        check(lib().de_ctx_set_variants(self.h, row, col))

    def set_profiling(self, on: bool):
        check(lib().de_ctx_set_profiling(self.h, int(on)))

    def stage_ms(self):
        ms = (C.c_double * 4)()
        runs = C.c_int()
        check(lib().de_ctx_stage_ms(self.h, ms, C.byref(runs)))
        return list(ms), runs.value


_default: Context | None = None


def default_context() -> Context:
    global _default
    if _default is None:
        _default = Context(0)
    return _default


def run_dense(L, R, cfg: Cfg | None = None, want_cand: bool = False):
    """(disp, margin, cand6 or None) as de_dense --threads 1 writes them."""
    return default_context().run_dense(L, R, cfg, want_cand)


def census(img, hw: int = 3, hh: int = 3) -> np.ndarray:
    return default_context().census(img, hw, hh)


def _close_default():
    global _default
    if _default is not None:
        _default.close()
        _default = None


import atexit as _atexit  # noqa: E402

_atexit.register(_close_default)
