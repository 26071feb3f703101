"""DEIR -> DEDD stream front end and the `de_dense_cuda` command line.

    python -m doubleeye_amd.stream LEFT.y8 RIGHT.y8 W H [--dmax N] [--iters N]
        [--min-margin F] [--sigma-s F] [--sigma-r F] [--ad F] [--ad-trunc N]
        [--lambda F] [--gamma F] [--damping F] [--threads N] [--out disp.f32]
        [--stream] [--no-confidence]

The same arguments as core/tools/de_dense_cuda.cu:402-460 (--threads is
accepted and ignored: the solve runs on the GPU). With --stream, DEIR packets
are read from stdin and DEDD packets written to stdout (de_dense_cuda.cu:44-59,
670-770); otherwise one pair is read from the two files and the disparity
written to --out as W*H float32. The loop is native (de_stream_run_fd); this
module parses arguments and offers packet helpers.
"""
from __future__ import annotations

import ctypes as C
import struct
import sys

import numpy as np

from ._lib import DeError, check, lib
from .dense import Cfg

__all__ = ["deir_packet", "parse_dedd", "Stream", "main"]

DEDD_CONF = 1


def deir_packet(img: np.ndarray, stream: int, frame: int, flags: int = 0) -> bytes:
    """One DEIR packet: "DEIR" | w u16 | h u16 | stream u8 | flags u8 | frame u32 | Y8."""
    img = np.ascontiguousarray(img, np.uint8)
    h, w = img.shape
    return b"DEIR" + struct.pack("<HHBBI", w, h, stream, flags, frame) + img.tobytes()


def parse_dedd(buf: bytes):
    """Split DEDD packets: list of dicts (frame, flags, left, q4, conf or None)."""
    out, pos = [], 0
    while pos + 16 <= len(buf):
        if buf[pos:pos + 4] != b"DEDD":
            raise ValueError("bad DEDD magic")
        w, h, frame, flags = struct.unpack_from("<HHII", buf, pos + 4)
        n = w * h
        pos += 16
        left = np.frombuffer(buf, np.uint8, n, pos).reshape(h, w)
        pos += n
        q4 = np.frombuffer(buf, np.int16, n, pos).reshape(h, w)
        pos += 2 * n
        conf = None
        if flags & DEDD_CONF:
            conf = np.frombuffer(buf, np.uint8, n, pos).reshape(h, w)
            pos += n
        out.append(dict(frame=frame, flags=flags, left=left, q4=q4, conf=conf))
    return out


class Stream:
    """Incremental DEIR -> DEDD over a context (de_stream_feed)."""

    def __init__(self, ctx, W: int, H: int, cfg: Cfg, want_conf: bool = True):
        self.W, self.H, self.conf = W, H, want_conf
        self._out = None
        h = C.c_void_p()
        check(lib().de_stream_create(ctx.h, W, H, C.byref(cfg.to_c()), int(want_conf),
                                     C.byref(h)))
        self.h = h.value

    def feed(self, data) -> bytes:
        """Feed DEIR bytes; return the DEDD packets of every pair they completed."""
        if not isinstance(data, bytes):
            data = bytes(data)
        # every complete image can finish at most one pair
        pairs = len(data) // (self.W * self.H) + 2
        cap = pairs * lib().de_stream_packet_bytes(self.W, self.H, int(self.conf))
        if self._out is None or self._out.size < cap:
            self._out = np.empty(max(cap, 1), np.uint8)      # kept for later calls
        n = C.c_size_t(0)
        check(lib().de_stream_feed(self.h, data, len(data), self._out.ctypes.data, cap,
                                   C.byref(n)))
        return self._out[: n.value].tobytes()

    def stats(self):
        p, d = C.c_long(0), C.c_long(0)
        check(lib().de_stream_stats(self.h, C.byref(p), C.byref(d)))
        return p.value, d.value

    def close(self):
        if self.h:
            lib().de_stream_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _parse(argv):
    if len(argv) < 4:
        sys.stderr.write(
            "usage: python -m doubleeye_amd.stream LEFT.y8 RIGHT.y8 W H [--dmax N (default 64)]\n"
            "          [--iters N] [--threads N] [--min-margin F] [--sigma-s F] [--sigma-r F]\n"
            "          [--ad F] [--ad-trunc N] [--lambda F] [--gamma F] [--damping F]\n"
            "          [--out disp.f32] [--stream] [--no-confidence]\n")
        return None
    lp, rp, W, H = argv[0], argv[1], int(argv[2]), int(argv[3])
    cfg = Cfg(dmax=64)
    opts = dict(stream=False, conf=True, out=None)
    floats = {"--min-margin": "min_margin", "--sigma-s": "sigma_s", "--sigma-r": "sigma_r",
              "--ad": "ad", "--lambda": "lam", "--gamma": "gamma", "--damping": "damping"}
    ints = {"--dmax": "dmax", "--iters": "iters", "--ad-trunc": "ad_trunc", "--threads": None}
    i = 4
    while i < len(argv):
        a = argv[i]
        has = i + 1 < len(argv)
        if a in floats and has:
            setattr(cfg, floats[a], float(argv[i + 1])); i += 2
        elif a in ints and has:
            if ints[a]:
                setattr(cfg, ints[a], int(argv[i + 1]))
            i += 2
        elif a == "--out" and has:
            opts["out"] = argv[i + 1]; i += 2
        elif a == "--stream":
            opts["stream"] = True; i += 1
        elif a == "--no-confidence":
            opts["conf"] = False; i += 1
        else:
            sys.stderr.write(f"unknown argument {a}\n")
            return None
    return lp, rp, W, H, cfg, opts


def main(argv=None) -> int:
    args = _parse(sys.argv[1:] if argv is None else argv)
    if args is None:
        return 2
    lp, rp, W, H, cfg, opts = args
    from .dense import Context
    ctx = Context(0)
    try:
        if opts["stream"]:
            sys.stdout.flush()
            pairs, dropped = C.c_long(0), C.c_long(0)
            rc = lib().de_stream_run_fd(ctx.h, sys.stdin.fileno(), sys.stdout.fileno(), W, H,
                                        C.byref(cfg.to_c()), int(opts["conf"]),
                                        C.byref(pairs), C.byref(dropped))
            sys.stderr.write(f"pairs {pairs.value}, dropped {dropped.value}\n")
            if rc != 0:
                sys.stderr.write(lib().de_last_error().decode() + "\n")
                return 1
            return 0
        try:
            L = np.fromfile(lp, np.uint8).reshape(H, W)
            R = np.fromfile(rp, np.uint8).reshape(H, W)
        except (OSError, ValueError):
            sys.stderr.write(f"failed to load {lp} / {rp} at {W}x{H}\n")
            return 1
        disp, _, _ = ctx.run_dense(L, R, cfg)
        if opts["out"]:
            disp.astype(np.float32).tofile(opts["out"])
        sys.stderr.write(f"coverage {100.0 * np.isfinite(disp).mean():.1f}%\n")
        return 0
    except DeError as e:
        sys.stderr.write(f"{e}\n")
        return 1
    finally:
        ctx.close()


if __name__ == "__main__":
    sys.exit(main())
