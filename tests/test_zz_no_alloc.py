"""de_dense_run_dev_noalloc (the allocation-free dense entry point) through the
C ABI against de_dense_run_dev on the same device inputs: disparity, margin
and the six candidate planes bit for bit, on two calls in a row (the second
smaller, reusing the static scratch the first one left), and a frame beyond
the scratch refused with an error."""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def note(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def run_both(ctx, L, R, cfg):
    H, W = L.shape
    dL, dR = ctx.to_device(L), ctx.to_device(R)
    outs = []
    for fn in (ctx.run_dense_dev, ctx.run_dense_dev_noalloc):
        d = ctx.empty((H, W), np.float32)
        m = ctx.empty((H, W), np.float32)
        c6 = ctx.empty((6, H, W), np.float32)
        ctx.fill_dev(d, 0xAB); ctx.fill_dev(m, 0xAB); ctx.fill_dev(c6, 0xAB)
        fn(dL.ptr, dR.ptr, W, H, cfg, d.ptr, m.ptr, c6.ptr)
        ctx.sync()
        outs.append((d.numpy(), m.numpy(), c6.numpy()))
        for a in (d, m, c6):
            a.free()
    dL.free(); dR.free()
    return outs


def test_noalloc_equals_run_dense_dev(ctx):
    from doubleeye_amd import synth
    from doubleeye_amd._lib import DeError, lib
    from doubleeye_amd.dense import Cfg
    assert lib().de_dense_noalloc_limit() >= 230 << 20
    cases = [
        ("848x480 D128 9x7 it8", synth.slanted_pair(848, 480, seed=2)[:2],
         Cfg(dmax=128, iters=8, sigma_s=20.0, sigma_r=0.0392, census_hw=4, census_hh=3)),
        ("300x160 D48 7x7 it3 lrc", synth.fronto_pair(300, 160, seed=9)[:2],
         Cfg(dmax=48, iters=3, sigma_s=20.0, sigma_r=0.0392, lrc=1)),
    ]
    for tag, (L, R), cfg in cases:
        note(f"[noalloc] {tag}")
        (d0, m0, c0), (d1, m1, c1) = run_both(ctx, L, R, cfg)
        for a, b, what in ((d0, d1, "disparity"), (m0, m1, "margin"), (c0, c1, "cand6")):
            assert np.array_equal(a.view(np.uint32), b.view(np.uint32)), f"{tag}: {what}"
        assert np.isfinite(d1).mean() > 0.5
    note("[noalloc] oversize frame refused")
    big = ctx.empty((720, 1280), np.uint8)
    out = ctx.empty((720, 1280), np.float32)
    with pytest.raises(DeError):
        ctx.run_dense_dev_noalloc(big.ptr, big.ptr, 1280, 720, Cfg(dmax=192), out.ptr, out.ptr)
    big.free(); out.free()
