"""Edge-list MASDA on the GPU (the NumPy call shapes) against the oracle
(oracle/masda_np.py) and the reference's own vectors: messages and histories
in float64 bit for bit; assignments equal to the oracle's with edge-index
tie-breaks, and to the reference's wherever no two beliefs are exactly equal."""
import sys

import numpy as np
import pytest

import masda_np_cases as mc

pytestmark = pytest.mark.gpu
SP, MX, DC = mc.load()


def note(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def test_masda_sparse(ctx):
    from doubleeye_amd import masda_stereo as gm
    from oracle import masda_np as mn
    for k, p in enumerate(SP):
        note(f"[edges] sparse {k} E={len(p['se'])}")
        a, hist = gm.masda_sparse(p["ei"], p["ej"], p["se"], p["m"], p["n"], ctx=ctx, **p["kw"])
        oa, ohist, bel = mn.masda_sparse(p["ei"], p["ej"], p["se"], p["m"], p["n"],
                                         return_belief=True, **p["kw"])
        assert np.array_equal(hist, ohist) and np.array_equal(hist, p["hist"]), k
        assert a == oa, k
        if not mc.has_belief_ties(bel):
            assert mc.as_set(mc.pairs(a)) == mc.as_set(p["assign"]), k


def test_masda_matrix(ctx):
    from doubleeye_amd import masda_stereo as gm
    from oracle import masda_np as mn
    for k, p in enumerate(MX):
        a, hist = gm.masda(p["S"], ctx=ctx, **p["kw"])
        oa, ohist = mn.masda(p["S"], **p["kw"])
        assert np.array_equal(hist, p["hist"]) and np.array_equal(hist, ohist), k
        assert a == oa and mc.as_set(mc.pairs(a)) == mc.as_set(p["assign"]), k


def test_decide(ctx):
    from doubleeye_amd import masda_stereo as gm
    for k, p in enumerate(DC):
        assert mc.as_set(mc.pairs(gm.decide(p["S"], p["bel"], p["lam"], ctx=ctx))) == \
            mc.as_set(p["assign"]), k


def test_seg_max_excluding(ctx):
    from doubleeye_amd import masda_stereo as gm
    from oracle import masda_np as mn
    rng = np.random.default_rng(2)
    for E, n in [(1, 1), (50, 7), (10000, 300), (200000, 5000)]:
        idx = rng.integers(0, n, E)
        vals = rng.integers(-4, 5, E).astype(np.float64) / 3.0
        assert np.array_equal(gm._seg_max_excluding(vals, idx, n, ctx=ctx),
                              mn.seg_max_excluding(vals, idx, n))


@pytest.mark.parametrize("shape", [(64, 24, 1, 12), (225, 188, 1, 30), (424, 240, 0, 48)])
def test_dense_problem_and_dense_masda(shape, ctx):
    from doubleeye_amd import masda_stereo as gm
    from doubleeye_amd import synth
    from oracle import masda_np as mn
    W, H, dmin, dmax = shape
    L, R, _ = synth.fronto_pair(W, H, 5)
    note(f"[edges] dense problem {W}x{H} d {dmin}..{dmax}")
    ei, ej, se, nodes = gm.dense_problem(L, R, dmin, dmax, ctx=ctx)
    oei, oej, ose, onodes = mn.dense_problem(L, R, dmin, dmax)
    assert nodes == onodes and np.array_equal(ei, oei) and np.array_equal(ej, oej)
    assert np.array_equal(se.view(np.uint32), ose.view(np.uint32))
    a, hist = gm.masda_sparse(ei, ej, se, nodes, nodes, -0.1, -0.1, iters=12, ctx=ctx)
    oa, ohist = mn.masda_sparse(oei, oej, ose, nodes, nodes, -0.1, -0.1, iters=12)
    assert np.array_equal(hist, ohist) and a == oa
    disp, dh = gm.dense_masda(L, R, dmin, dmax, iters=12, ctx=ctx)
    want = np.full((H, W), np.nan, np.float32)
    for i, j in oa.items():
        y, x = divmod(i, W)
        want[y, x] = x - (j % W)
    assert np.array_equal(dh, ohist)
    assert np.array_equal(np.isnan(disp), np.isnan(want))
    assert np.array_equal(disp[~np.isnan(disp)], want[~np.isnan(want)])


def test_out_of_range_indices_raise(ctx):
    from doubleeye_amd import masda_stereo as gm
    from doubleeye_amd._lib import DeError
    with pytest.raises(DeError):
        gm.masda_sparse(np.array([0, 5]), np.array([0, 0]), np.array([0.5, 0.5]), 3, 3, ctx=ctx)
