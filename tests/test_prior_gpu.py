"""The prior-masked dense sweep on the GPU (de_dense_run_prior, de_dense_run_c2f)
against the oracle and the reference's own `de_dense --prior` / `--c2f` output
(tests/golden/make_prior_golden.py). Bar as for the full sweep: candidates,
margins and planes bit-exact; disparities bit-exact except at the oracle's
exact-tie pixels. --c2f's coarse level is a full sweep at half resolution, so
it too may differ at its exact-tie pixels; the fine sweep is then checked
against the oracle's masked sweep under the GPU's own coarse map, and against
the reference where the coarse level has no ties."""
import sys

import numpy as np
import pytest

import prior_cases as pc
from test_dense_gpu import dcfg, same_f32

pytestmark = pytest.mark.gpu


def note(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def compare(tag, d, m, cand, od, om, oc, ties):
    assert same_f32(m, om).all(), f"{tag}: margin differs at {(~same_f32(m, om)).sum()} px"
    for p in range(6):
        eq = same_f32(cand[p], oc[p])
        if p in (4, 5):
            eq |= ties.astype(bool)
        assert eq.all(), f"{tag}: cand plane {p} differs at {(~eq).sum()} px"
    bad = ~same_f32(d, od) & ~ties.astype(bool)
    assert not bad.any(), f"{tag}: {bad.sum()} non-tie pixels differ (ties {ties.sum()})"
    assert np.isfinite(d).sum() == np.isfinite(od).sum()


@pytest.mark.parametrize("name", pc.cases())
def test_prior_and_c2f_equal_reference(name, ctx, oracle):
    L, R, kw, band, c2f, prior, d_ref, c_sha = pc.case(name)
    note(f"[prior] {name} {L.shape} band {band} c2f {c2f}")
    if c2f:
        d, m, cand, coarse, used = ctx.run_dense_c2f(L, R, dcfg(**kw), band, want_cand=True)
        coarse_ties = check_coarse(L, R, kw, coarse, oracle)
        assert used
        od, om, oc, ties = oracle.dense_run_prior(L, R, oracle.cfg(**kw), coarse, band,
                                                  want_cand=True, want_ties=True)
        if coarse_ties:
            compare(name, d, m, cand, od, om, oc, ties)
            return
    else:
        d, m, cand = ctx.run_dense_prior(L, R, prior, dcfg(**kw), band, want_cand=True)
        od, om, oc, ties = oracle.dense_run_prior(L, R, oracle.cfg(**kw), prior, band,
                                                  want_cand=True, want_ties=True)
    compare(name, d, m, cand, od, om, oc, ties)
    eq = same_f32(d, d_ref) | ties.astype(bool)
    assert eq.all(), f"{name}: {(~eq).sum()} px differ from the reference"
    if not ties.any():
        assert pc.sha(cand) == c_sha


def check_coarse(L, R, kw, coarse, oracle):
    """The coarse level against the oracle's half-resolution sweep: equal except
    at that sweep's exact-tie pixels and the holes filled from them. Returns
    whether the coarse sweep had ties."""
    def half(a):
        a = a.astype(np.int32)
        H2, W2 = a.shape[0] // 2, a.shape[1] // 2
        a = a[: 2 * H2, : 2 * W2]
        return ((a[0::2, 0::2] + a[0::2, 1::2] + a[1::2, 0::2] + a[1::2, 1::2] + 2) // 4
                ).astype(np.uint8)
    ck = dict(kw, dmin=max(1, kw.get("dmin", 0) // 2), dmax=(kw.get("dmax", 64) + 1) // 2,
              min_margin=0.0, iters=1, subpixel=0)
    raw, _, _, ties = oracle.dense_run(half(L), half(R), oracle.cfg(**ck), want_ties=True)
    ok = same_f32(coarse, raw) | ties.astype(bool) | ~np.isfinite(raw)
    assert ok.all(), f"coarse level differs at {(~ok).sum()} non-tie pixels"
    if not ties.any():
        assert same_f32(coarse[np.isfinite(raw)], raw[np.isfinite(raw)]).all()
    return bool(ties.any())


@pytest.mark.parametrize("seed,band,kw", [
    (31, 0, dict(dmax=48)),
    (32, 3, dict(dmax=96, iters=3, lrc=1)),
    (33, 1, dict(dmin=5, dmax=70, census_hw=4)),
])
def test_random_priors_equal_oracle(seed, band, kw, ctx, oracle):
    """Priors with jumps, out-of-range values and a disparity range not
    starting at 0; one plane count above 64 (two groups)."""
    from doubleeye_amd import synth
    L, R, gt = synth.fronto_pair(300, 160, seed=seed)
    rng = np.random.default_rng(seed)
    prior = np.where(rng.random(gt.shape) < 0.1, rng.uniform(-20, 140, gt.shape),
                     np.nan_to_num(gt, nan=20.0) + rng.normal(0, 2, gt.shape)).astype(np.float32)
    note(f"[prior] random seed {seed} band {band} {kw}")
    d, m, cand = ctx.run_dense_prior(L, R, prior, dcfg(**kw), band, want_cand=True)
    od, om, oc, ties = oracle.dense_run_prior(L, R, oracle.cfg(**kw), prior, band,
                                              want_cand=True, want_ties=True)
    compare(f"random {seed}", d, m, cand, od, om, oc, ties)


def test_empty_mask_and_errors(ctx, oracle):
    """A prior outside the disparity range admits no plane: no candidates at
    all. Holes in a prior and a wrong prior size are refused."""
    from doubleeye_amd import synth
    from doubleeye_amd._lib import DeError
    L, R, _ = synth.fronto_pair(120, 64, seed=3)
    prior = np.full(L.shape, 500.0, np.float32)
    d, m, cand = ctx.run_dense_prior(L, R, prior, dcfg(dmax=32), 2, want_cand=True)
    assert not np.isfinite(d).any()
    od, _, _, _ = oracle.dense_run_prior(L, R, oracle.cfg(dmax=32), prior, 2)
    assert not np.isfinite(od).any()
    bad = prior.copy()
    bad[3, 4] = np.nan
    with pytest.raises(DeError):
        ctx.run_dense_prior(L, R, bad, dcfg(dmax=32))
    with pytest.raises(DeError):
        ctx.run_dense_prior(L, R, prior[:10], dcfg(dmax=32))


@pytest.mark.parametrize("shape,kw", [((480, 848), dict(dmax=127, iters=8, subpixel=0)),
                                      ((121, 203), dict(dmax=40, iters=3, subpixel=0))])
def test_c2f_device_pointers_equal_host_call(ctx, shape, kw):
    """de_dense_run_c2f_dev on resident buffers gives the host call's outputs,
    bit for bit, with and without the candidate planes."""
    from doubleeye_amd import synth
    H, W = shape
    L, R, _ = synth.slanted_pair(W, H, seed=2, dhi=0.8 * kw["dmax"])[:3]
    note(f"[c2f dev] {W}x{H} {kw}")
    cfg = dcfg(**kw)
    d, m, cand, _, used = ctx.run_dense_c2f(L, R, cfg, 2, want_cand=True)
    dL, dR = ctx.to_device(L), ctx.to_device(R)
    dd, dm = ctx.empty((H, W), np.float32), ctx.empty((H, W), np.float32)
    dc = ctx.empty((6, H, W), np.float32)
    used_dev = ctx.run_dense_c2f_dev(dL.ptr, dR.ptr, W, H, cfg, dd.ptr, dm.ptr, dc.ptr, 2)
    ctx.sync()
    assert used_dev == used
    assert same_f32(dd.numpy(), d).all() and same_f32(dm.numpy(), m).all()
    assert same_f32(dc.numpy(), cand).all()
    used2 = ctx.run_dense_c2f_dev(dL.ptr, dR.ptr, W, H, cfg, dd.ptr, dm.ptr, 0, 2)
    ctx.sync()
    assert used2 == used and same_f32(dd.numpy(), d).all()
