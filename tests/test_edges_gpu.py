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


def test_census_hamming_build_problem_match_oracle(ctx):
    """census(img, pts, ...), hamming(a, b) and build_problem(...) on the
    device equal the oracle's restatement of masda_stereo.py:264-326 bit for
    bit: descriptors, distances, and every score (float32 and float64
    points, parameters on the gate boundaries)."""
    from doubleeye_amd import masda_stereo as gm
    from doubleeye_amd import synth
    from oracle import masda_np as mn
    L, R, _ = synth.fronto_pair(300, 120, seed=31)
    rng = np.random.default_rng(31)
    for hw, hh in ((3, 3), (4, 3), (2, 1)):
        pts = np.stack([rng.uniform(hw, 300 - hw, 500), rng.uniform(hh, 120 - hh, 500)], 1)
        pts = pts.astype(np.float32)
        pts[:4] = [[0, 0], [-3, -2], [299 - hw, 119 - hh], [hw, hh]]       # numpy wrap-around
        note(f"[np-shapes] census {2*hw+1}x{2*hh+1}")
        got = gm.census(L, pts, hw, hh, ctx=ctx)
        assert got.dtype == np.uint64
        assert np.array_equal(got, mn.census_points(L, pts, hw, hh))
    with pytest.raises(IndexError):
        gm.census(L, np.array([[297.0, 5.0]], np.float32), 3, 3, ctx=ctx)
    a = rng.integers(0, 2 ** 63, (70, 1), dtype=np.uint64)
    b = rng.integers(0, 2 ** 63, (1, 90), dtype=np.uint64)
    h = gm.hamming(a, b, ctx=ctx)
    assert h.dtype == np.int32 and h.shape == (70, 90)
    assert np.array_equal(h, mn.hamming(a, b))
    # keypoint-like problem: left/right points on the same rows, descriptors
    # from the census, plus points on the gate boundaries
    m, n = 700, 650
    pl = np.stack([rng.uniform(10, 290, m), rng.integers(5, 115, m) + rng.choice([0, .25, .5], m)], 1)
    pr = np.stack([rng.uniform(10, 290, n), rng.integers(5, 115, n) + rng.choice([0, .25, .5], n)], 1)
    pr[:40] = pl[:40] - [[1.0, 1.5]] * 40                                # d = 1, dy = 1.5
    pr[40:80] = pl[40:80] - [[60.0, -1.5]] * 40                          # d = 60, dy = -1.5
    pr[80:120] = pl[80:120] - [[60.5, 0.0]] * 40                         # just outside
    for t in (np.float32, np.float64):
        plt, prt = pl.astype(t), pr.astype(t)
        dl = mn.census_points(L, plt, 3, 3)
        dr = mn.census_points(R, prt, 3, 3)
        for kw in (dict(), dict(max_dy=0.5, dmin=2.0, dmax=40.0, sigma_y=0.7, w_y=0.3),
                   dict(sigma_y=3.0, w_y=1.7)):
            note(f"[np-shapes] build_problem {t.__name__} {kw}")
            got = gm.build_problem(plt, dl, prt, dr, ctx=ctx, **kw)
            want = mn.build_problem(plt, dl, prt, dr, **kw)
            assert got.dtype == np.float32 and got.shape == (m, n)
            assert np.array_equal(got.view(np.uint32), want.view(np.uint32)), (t, kw)
            assert np.isfinite(got).sum() > 100
