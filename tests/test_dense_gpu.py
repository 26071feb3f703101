"""Parity of the HIP dense path (through the C ABI) with the CPU oracle.

Bar: census bits, filtered int16 planes, top-2 candidates (values, indices,
sub-pixel neighbours), margins and FP32 messages bit-exact; disparities
bit-exact except at documented exact-tie pixels (two claimants of one right
pixel with exactly equal belief, ordered by std::sort's unstable partition in
the reference). Coverage is tie-invariant and must match exactly.
"""
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
    """Bitwise equality with every NaN treated as equal."""
    a, b = np.asarray(a, np.float32), np.asarray(b, np.float32)
    both_nan = np.isnan(a) & np.isnan(b)
    return (a.view(np.uint32) == b.view(np.uint32)) | both_nan


def dcfg(**kw):
    from doubleeye_amd.dense import Cfg
    if "lambda_" in kw:
        kw["lam"] = kw.pop("lambda_")
    return Cfg(**kw)


def check_dense(ctx, oracle, L, R, kw, tag):
    note(f"[dense] {tag} {L.shape} {kw}")
    d, m, cand = ctx.run_dense(L, R, dcfg(**kw), want_cand=True)
    od, om, oc, ties = oracle.dense_run(L, R, oracle.cfg(**kw), want_cand=True, want_ties=True)
    assert same_f32(m, om).all(), f"{tag}: margin differs at {(~same_f32(m, om)).sum()} px"
    for p in (0, 1, 2, 3, 4, 5):
        eq = same_f32(cand[p], oc[p])
        if p in (4, 5):   # lrc, lrc2 follow the disparity; only tie pixels may move them
            eq |= ties.astype(bool)
        assert eq.all(), f"{tag}: cand plane {p} differs at {(~eq).sum()} px"
    eq = same_f32(d, od)
    bad = ~eq & ~ties.astype(bool)
    # a tie swap moves an answer between two tied claimants: the loser of one
    # side is the winner of the other, so the claimants' partner pixels may
    # differ too; every differing pixel must be a flagged tie pixel
    assert not bad.any(), f"{tag}: {bad.sum()} non-tie pixels differ (ties {ties.sum()})"
    assert np.isfinite(d).sum() == np.isfinite(od).sum(), "coverage must be tie-invariant"
    return d, od, ties


def test_census_matches_oracle(ctx, oracle):
    rng = np.random.default_rng(11)
    for (h, w) in ((37, 91), (64, 64), (9, 7), (200, 133)):
        img = rng.integers(0, 256, (h, w), dtype=np.uint8)
        for hw, hh in ((3, 3), (4, 3), (2, 2), (1, 1)):
            note(f"[census] {h}x{w} window {2*hw+1}x{2*hh+1}")
            got = ctx.census(img, hw, hh)
            np.testing.assert_array_equal(got, oracle.census(img, hw, hh))


@pytest.mark.parametrize("win", [(3, 3), (4, 3)])
def test_filtered_volume_bit_exact(ctx, oracle, win):
    from doubleeye_amd import synth
    L, R, _ = synth.fronto_pair(150, 70, seed=5, disparities=(3, 9, 17))
    kw = dict(dmax=40, sigma_s=20.0, sigma_r=0.0392, census_hw=win[0], census_hh=win[1])
    note(f"[volume] 150x70 D=40 window {win}")
    got = ctx.volume(L, R, dcfg(**kw), 0, 40)
    want = oracle.filtered_planes(L, R, oracle.cfg(**kw), 0, 40)
    np.testing.assert_array_equal(got, want)


def test_c1_fixture_all_flag_sets(ctx, oracle):
    z = np.load(os.path.join(GOLDEN, "dense_c1.npz"))
    for name in ("c1", "default48", "nosub_margin", "parabola_ratio", "d96_it8", "lrc_it2"):
        kw = flags_to_kwargs(str(z[f"{name}_flags"]))
        d, od, ties = check_dense(ctx, oracle, z["L"], z["R"], kw, name)
        if name == "c1":
            # the oracle equals the reference here (pinned by test_oracle)
            eq = same_f32(d, z["c1_disp"]) | ties.astype(bool)
            assert eq.all()


def test_exact_reverse_match_plane(ctx, oracle):
    """--lrc (de_dense.cpp:1438-1453, 1581-1597) on larger, slanted pairs and
    the 9x7 window; plane 4 is -1e30 everywhere without --lrc."""
    from doubleeye_amd import synth
    for (W, H, D, win, it) in ((424, 240, 96, (3, 3), 4), (320, 200, 64, (4, 3), 0)):
        L, R, _ = synth.slanted_pair(W, H, seed=12, dhi=0.7 * D)
        kw = dict(dmax=D, iters=it, sigma_s=20.0, sigma_r=0.0392, lrc=1,
                  census_hw=win[0], census_hh=win[1])
        check_dense(ctx, oracle, L, R, kw, f"lrc {W}x{H} D{D} {win}")
    kw["lrc"] = 0
    _, _, cand = ctx.run_dense(L, R, dcfg(**kw), want_cand=True)
    assert (cand[4] == np.float32(-1e30)).all()


@pytest.mark.parametrize("scene", ["night_off", "morning_on", "evening_on"])
def test_scene_coverage_known_answer(ctx, oracle, scene):
    z = np.load(os.path.join(GOLDEN, "scenes.npz"))
    d, od, ties = check_dense(ctx, oracle, z[f"{scene}_L"], z[f"{scene}_R"], dict(dmax=64), scene)
    cov = round(100.0 * np.isfinite(d).mean(), 1)
    assert cov == float(z[f"{scene}_coverage_pct"])


def test_c2_full_size(ctx, oracle):
    from doubleeye_amd import synth
    L, R, _ = synth.slanted_pair(848, 480, seed=2)
    kw = dict(dmax=128, iters=8, sigma_s=20.0, sigma_r=0.0392)
    check_dense(ctx, oracle, L, R, kw, "C2 848x480x128")


def test_c2_full_size_census_9x7(ctx, oracle):
    """C2 as stated: 848x480, 128 disparities, 8 iterations, census 9x7
    (62 bits, score level (31 - h)/31, x-border 4)."""
    from doubleeye_amd import synth
    L, R, _ = synth.slanted_pair(848, 480, seed=2)
    kw = dict(dmax=128, iters=8, sigma_s=20.0, sigma_r=0.0392, census_hw=4, census_hh=3)
    check_dense(ctx, oracle, L, R, kw, "C2 848x480x128 census 9x7")


def test_c1_census_9x7(ctx, oracle):
    """C1 as stated: 424x240 random-dot pair on fronto planes at 8/20/35 px,
    48 disparities, 5 iterations, census 9x7."""
    from doubleeye_amd import synth
    L, R, _ = synth.fronto_pair(424, 240, seed=1, disparities=(8, 20, 35))
    kw = dict(dmax=48, iters=5, sigma_s=20.0, sigma_r=0.0392, census_hw=4, census_hh=3)
    d, od, ties = check_dense(ctx, oracle, L, R, kw, "C1 424x240x48 census 9x7")
    assert np.isfinite(d).mean() > 0.5


@pytest.mark.parametrize("w,h,d", [(424, 560, 128), (1280, 720, 192)])
def test_multi_round_row_pass(ctx, oracle, w, h, d):
    """More row waves than SIMDs (560 x 2 and 720 x 3 against 1,024): the
    register-resident row pass runs in rounds, without LDS forward slots; the
    first frame's column pass keeps six register blocks and the LDS block."""
    from doubleeye_amd import synth
    L, R, _ = synth.slanted_pair(w, h, seed=5)
    kw = dict(dmax=d, iters=4, sigma_s=20.0, sigma_r=0.0392, census_hw=4, census_hh=3)
    check_dense(ctx, oracle, L, R, kw, f"multi-round {w}x{h}x{d}")


def test_9x7_extension_and_dmin0(ctx, oracle):
    from doubleeye_amd import synth
    L, R, _ = synth.fronto_pair(300, 160, seed=9)
    check_dense(ctx, oracle, L, R, dict(dmax=47, dmin=0, census_hw=4, iters=3), "9x7 dmin0")


@pytest.mark.parametrize("shape,kw", [
    ((240, 424), dict(dmax=192, iters=2)),        # G = 3, most pixels have few planes
    ((96, 130), dict(dmax=256, dmin=1)),          # D = 256, more than the width allows
    ((140, 300), dict(dmax=250, iters=2)),        # G = 4, two register blocks in the column pass
    ((200, 260), dict(dmax=250, iters=2)),        # G = 4, three register blocks
    ((20, 300), dict(dmax=64, iters=4)),          # few rows
    ((7, 7), dict(dmax=4)),                       # one valid pixel at most
    ((6, 40), dict(dmax=8)),                      # no valid row
    ((50, 70), dict(dmax=1, iters=3)),            # a single disparity
    ((33, 65), dict(dmax=3, min_margin=0.02, fit_eq=0)),
    ((30, 1100), dict(dmax=70, iters=3)),         # rows wider than the solver's threads
    ((24, 2500), dict(dmax=40, iters=3)),         # several pixels per solver thread
    ((10, 4096), dict(dmax=24, iters=2)),         # the widest row supported
])
def test_edge_shapes(ctx, oracle, shape, kw):
    rng = np.random.default_rng(sum(shape))
    L = rng.integers(0, 256, shape, dtype=np.uint8)
    R = np.roll(L, -2, axis=1) // 2 + rng.integers(0, 60, shape, dtype=np.uint8)
    check_dense(ctx, oracle, L, R, kw, f"edge {shape}")


def test_invalid_arguments_raise(ctx):
    from doubleeye_amd._lib import DeError
    L = np.zeros((32, 32), np.uint8)
    for kw in (dict(dmax=0), dict(dmax=400), dict(census_hw=5), dict(sigma_s=0.0),
               dict(iters=-1)):
        with pytest.raises(DeError):
            ctx.run_dense(L, L, dcfg(**kw))
    with pytest.raises(ValueError):
        ctx.run_dense(L, np.zeros((32, 31), np.uint8))
    wide = np.zeros((8, 4097), np.uint8)          # one column beyond the row solver
    with pytest.raises(DeError):
        ctx.run_dense(wide, wide, dcfg(dmax=16))
