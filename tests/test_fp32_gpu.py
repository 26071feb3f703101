"""The FP32 full-volume path on the GPU: the float volume equals the
reference's `de_dense --dump-vol` bit for bit, and `--topk 0` disparities and
margins equal the reference's / the oracle's (disparities except at documented
exact-tie pixels)."""
import os
import sys

import numpy as np
import pytest

from conftest import GOLDEN, flags_to_kwargs

pytestmark = pytest.mark.gpu


def note(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def same_f32(a, b):
    return (a.view(np.uint32) == b.view(np.uint32)) | (np.isnan(a) & np.isnan(b))


def test_volume_equals_reference_dump(ctx):
    from doubleeye_amd.dense import Cfg
    z = np.load(os.path.join(GOLDEN, "volume_f32.npz"))
    vol = ctx.volume_f32(z["L"], z["R"], Cfg(**flags_to_kwargs(str(z["flags"]))))
    assert np.array_equal(vol.view(np.uint32), z["vol"].view(np.uint32))


@pytest.mark.parametrize("shape", [(424, 240, 1, 48, 3, 3), (200, 120, 0, 64, 4, 3),
                                   (96, 40, 5, 40, 3, 3), (640, 360, 1, 200, 4, 3)])
def test_volume_equals_oracle(shape, ctx, oracle):
    from doubleeye_amd import synth
    from doubleeye_amd.dense import Cfg
    W, H, dmin, dmax, hw, hh = shape
    L, R, _ = synth.slanted_pair(W, H, seed=6, dhi=min(0.8 * dmax, 100.0))
    kw = dict(dmin=dmin, dmax=dmax, census_hw=hw, census_hh=hh, sigma_s=20.0, sigma_r=0.0392)
    note(f"[fp32] volume {shape}")
    got = ctx.volume_f32(L, R, Cfg(**kw))
    want = oracle.volume_f32(L, R, oracle.cfg(**kw))
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))


FULL = np.load(os.path.join(GOLDEN, "dense_full.npz"))
FULL_CASES = [(p, n) for p in ("fronto", "slanted") for n in ("it4", "it0", "margin", "nosub")]


@pytest.mark.parametrize("case", FULL_CASES, ids=["-".join(c) for c in FULL_CASES])
def test_full_solver_equals_reference(case, ctx, oracle):
    from doubleeye_amd.dense import Cfg
    p, n = case
    kw = flags_to_kwargs(str(FULL[f"{p}_{n}_flags"]).replace("--topk 0", ""))
    L, R = FULL[f"{p}_L"], FULL[f"{p}_R"]
    note(f"[fp32] topk0 {p} {n}")
    d, m = ctx.run_dense_full(L, R, Cfg(**kw))
    od, om, ties = oracle.dense_run_full(L, R, oracle.cfg(**kw), want_ties=True)
    assert np.array_equal(m.view(np.uint32), om.view(np.uint32))
    assert (same_f32(d, od) | (ties > 0)).all()
    assert (same_f32(d, FULL[f"{p}_{n}_disp"]) | (ties > 0)).all()


@pytest.mark.parametrize("shape", [(320, 240, 0, 96, 6), (848, 480, 1, 128, 8)])
def test_full_solver_equals_oracle(shape, ctx, oracle):
    from doubleeye_amd import synth
    from doubleeye_amd.dense import Cfg
    W, H, dmin, dmax, it = shape
    L, R, _ = synth.slanted_pair(W, H, seed=8, dhi=min(0.8 * dmax, 110.0))
    kw = dict(dmin=dmin, dmax=dmax, iters=it, sigma_s=20.0, sigma_r=0.0392)
    note(f"[fp32] topk0 {shape}")
    d, m = ctx.run_dense_full(L, R, Cfg(**kw))
    od, om, ties = oracle.dense_run_full(L, R, oracle.cfg(**kw), want_ties=True)
    assert np.array_equal(m.view(np.uint32), om.view(np.uint32))
    assert (same_f32(d, od) | (ties > 0)).all()
    assert np.isfinite(d).sum() == np.isfinite(od).sum() or ties.any()
