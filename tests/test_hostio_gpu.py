"""Host-pointer entry points through the context's pinned staging and through
de_host_alloc memory: every output equals the resident call's bit for bit,
whatever the mix of pageable and pinned buffers, across the 1 MiB chunks the
outputs come back in."""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def note(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def resident(ctx, L, R, cfg):
    H, W = L.shape
    dL, dR = ctx.to_device(L), ctx.to_device(R)
    dd, dm = ctx.empty((H, W), np.float32), ctx.empty((H, W), np.float32)
    dc = ctx.empty((6, H, W), np.float32)
    ctx.run_dense_dev(dL.ptr, dR.ptr, W, H, cfg, dd.ptr, dm.ptr, dc.ptr)
    ctx.sync()
    out = dd.numpy(), dm.numpy(), dc.numpy()
    for b in (dL, dR, dd, dm, dc):
        b.free()
    return out


def same(a, b):
    return np.array_equal(np.asarray(a).view(np.uint32), np.asarray(b).view(np.uint32))


@pytest.mark.parametrize("shape,dmax", [((480, 848), 128), ((97, 203), 40)])
def test_host_run_pageable_and_pinned(ctx, shape, dmax):
    from doubleeye_amd import synth
    from doubleeye_amd.dense import Cfg, host_empty
    H, W = shape
    L, R, _ = synth.slanted_pair(W, H, seed=5)
    cfg = Cfg(dmax=dmax, iters=4, sigma_s=20.0, sigma_r=0.0392)
    note(f"[hostio] {W}x{H} D={dmax}")
    want = resident(ctx, L, R, cfg)
    # pageable in and out (cand6 spans several chunks at 848x480)
    got = ctx.run_dense(L, R, cfg, want_cand=True)
    assert all(same(g, w) for g, w in zip(got, want))
    # pinned outputs, pageable inputs; twice, so the staging is reused
    out = (host_empty((H, W), np.float32), host_empty((H, W), np.float32),
           host_empty((6, H, W), np.float32))
    for _ in range(2):
        for o in out:
            o.fill(np.nan)
        got = ctx.run_dense(L, R, cfg, want_cand=True, out=out)
        assert all(g is o for g, o in zip(got, out))
        assert all(same(g, w) for g, w in zip(got, want))
    # pinned inputs (interior views of one allocation), pageable outputs
    big = host_empty((2 * H + 3, W), np.uint8)
    pL, pR = big[1:H + 1], big[H + 2:2 * H + 2]
    pL[...] = L
    pR[...] = R
    got = ctx.run_dense(pL, pR, cfg)
    assert same(got[0], want[0]) and same(got[1], want[1])


def test_prior_and_c2f_host_calls_match_resident(ctx):
    from doubleeye_amd import synth
    from doubleeye_amd.dense import Cfg
    L, R, gt = synth.slanted_pair(424, 240, seed=7)[:3]
    H, W = L.shape
    cfg = Cfg(dmax=63, iters=3, subpixel=0)
    prior = np.nan_to_num(gt, nan=float(np.nanmedian(gt))).astype(np.float32)
    note("[hostio] c2f and prior, 424x240")
    d, m, _, coarse, used = ctx.run_dense_c2f(L, R, cfg, 2)
    dL, dR = ctx.to_device(L), ctx.to_device(R)
    dd, dm = ctx.empty((H, W), np.float32), ctx.empty((H, W), np.float32)
    ctx.run_dense_c2f_dev(dL.ptr, dR.ptr, W, H, cfg, dd.ptr, dm.ptr, 0, 2)
    ctx.sync()
    assert same(d, dd.numpy()) and same(m, dm.numpy())
    assert coarse.shape == (H // 2, W // 2) and np.isfinite(coarse).any()
    d1, m1, _ = ctx.run_dense_prior(L, R, prior, cfg, 2)
    d2, m2, _ = ctx.run_dense_prior(L, R, prior, cfg, 2)
    assert same(d1, d2) and same(m1, m2)
    for b in (dL, dR, dd, dm):
        b.free()

