"""The DEDD output planes of the HIP solver -- Q4 disparity and the u8
confidence byte (core/tools/de_dense_cuda.cu:731-745 over
core/include/doubleeye/dense_solve.hpp:158-171) -- against the oracle.

Q4 is integer work on the disparity: bit-exact. The confidence byte rounds
P * 255, where P comes from logf/expf: the device's and glibc's differ by an
ulp or so on some arguments, so a byte may differ by one where the oracle's
P * 255 lies next to a .5 boundary. The test allows exactly that, bounds it,
and reports the count:
  * every differing byte differs by 1;
  * the oracle's P * 255 there lies within 2e-3 of k + 0.5;
  * at most 1e-3 of the answered pixels differ.
The oracle's P * 255 is computed from the same (s1, s2, lrc) floats the GPU
used (the GPU's own --out-conf planes, themselves bit-exact against the
oracle), so the comparison isolates the confidence function; a second check
runs end to end from the oracle's planes, away from exact-tie pixels.
"""
import os
import sys

import numpy as np
import pytest

from conftest import GOLDEN, check_conf_bytes, flags_to_kwargs

pytestmark = pytest.mark.gpu


def note(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def run_planes(ctx, L, R, cfg):
    """One pair through the batch entry point with every output plane."""
    from doubleeye_amd.dense import Pair
    H, W = L.shape
    bufs = dict(L=ctx.to_device(np.ascontiguousarray(L)), R=ctx.to_device(np.ascontiguousarray(R)),
                disp=ctx.empty((H, W), np.float32), margin=ctx.empty((H, W), np.float32),
                cand=ctx.empty((6, H, W), np.float32), conf=ctx.empty((H, W), np.uint8),
                q4=ctx.empty((H, W), np.int16))
    try:
        p = Pair(bufs["L"].ptr, bufs["R"].ptr, W, H, bufs["disp"].ptr, bufs["margin"].ptr,
                 bufs["cand"].ptr, bufs["conf"].ptr, bufs["q4"].ptr)
        ctx.dense_batch_dev([p], [cfg])
        ctx.sync()
        return (bufs["disp"].numpy(), bufs["cand"].numpy(), bufs["conf"].numpy(),
                bufs["q4"].numpy())
    finally:
        for b in bufs.values():
            b.free()


def check_case(ctx, oracle, L, R, kw, tag):
    from doubleeye_amd.dense import Cfg
    lam = kw.get("lam", -0.1)
    d, c6, conf, q4 = run_planes(ctx, L, R, Cfg(**kw))
    # the confidence function on the GPU's own planes
    oq4, oconf, p255 = oracle.dedd_planes(d, c6, lam)
    np.testing.assert_array_equal(q4, oq4, err_msg=f"{tag}: Q4 plane")
    answered = q4 != -32768
    assert (conf[~answered] == 0).all()
    n = check_conf_bytes(conf, oconf, p255, answered, tag)
    # end to end from the oracle's own planes, away from exact-tie pixels
    od, _, oc, ties = oracle.dense_run(L, R, oracle.cfg(**kw), want_cand=True, want_ties=True)
    eq4, econf, ep255 = oracle.dedd_planes(od, oc, lam)
    keep = ~ties.astype(bool)
    np.testing.assert_array_equal(q4[keep], eq4[keep], err_msg=f"{tag}: Q4 vs oracle")
    check_conf_bytes(np.where(keep, conf, 0), np.where(keep, econf, 0),
                     np.where(keep, ep255, 0), answered & keep, tag + " (end to end)")
    return n


def test_c1_fixture_flag_sets(ctx, oracle):
    z = np.load(os.path.join(GOLDEN, "dense_c1.npz"))
    for name in ("c1", "default48", "nosub_margin", "parabola_ratio", "d96_it8", "lrc_it2"):
        kw = flags_to_kwargs(str(z[f"{name}_flags"]))
        check_case(ctx, oracle, z["L"], z["R"], kw, name)


@pytest.mark.parametrize("win", [(3, 3), (4, 3)])
def test_c2_full_size(ctx, oracle, win):
    from doubleeye_amd import synth
    L, R, _ = synth.slanted_pair(848, 480, seed=2)
    kw = dict(dmax=128, iters=8, sigma_s=20.0, sigma_r=0.0392, census_hw=win[0],
              census_hh=win[1])
    check_case(ctx, oracle, L, R, kw, f"C2 848x480x128 census {2*win[0]+1}x{2*win[1]+1}")


@pytest.mark.parametrize("scene", ["night_off", "morning_on", "evening_on"])
def test_ir_scenes(ctx, oracle, scene):
    z = np.load(os.path.join(GOLDEN, "scenes.npz"))
    check_case(ctx, oracle, z[f"{scene}_L"], z[f"{scene}_R"], dict(dmax=64, iters=2), scene)
