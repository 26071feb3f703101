"""The oracle's prior-masked sweep (`de_dense --prior`) and coarse-to-fine
driver (`--c2f`) against the reference binary's own output on the fixture of
tests/golden/make_prior_golden.py, plus the mask's closed-form properties.
CPU only."""
import numpy as np
import pytest

import prior_cases as pc


@pytest.mark.parametrize("name", pc.cases())
def test_oracle_equals_reference(name, oracle):
    L, R, kw, band, c2f, prior, d_ref, c_sha = pc.case(name)
    c = oracle.cfg(**kw)
    if c2f:
        d, m, cand, _, coarse, used = oracle.dense_run_c2f(L, R, c, band, want_cand=True)
        assert used and np.isfinite(coarse).all()
    else:
        d, m, cand, _ = oracle.dense_run_prior(L, R, c, prior, band, want_cand=True)
    assert np.array_equal(d.view(np.uint32), d_ref.view(np.uint32))
    assert pc.sha(cand) == c_sha


def test_wide_band_equals_nothing_filtered_outside(oracle):
    """Closed form: a band wider than the disparity range with a constant prior
    admits every plane at every pixel, so every rectangle is the whole score
    region; the masked filter then starts at the region's edge instead of
    the image's, which is the only difference from the full sweep: interior
    candidates far from the edges agree to within the filter's decay."""
    from doubleeye_amd import synth
    L, R, _ = synth.fronto_pair(160, 96, seed=7)
    c = oracle.cfg(dmax=24, subpixel=0)
    full = oracle.dense_run(L, R, c, want_cand=True)[2]
    prior = np.full(L.shape, 12.0, np.float32)
    masked = oracle.dense_run_prior(L, R, c, prior, 100, want_cand=True)[2]
    # the best candidate's disparity agrees away from the borders
    core = (slice(24, -24), slice(60, -24))
    agree = (full[2][core] == masked[2][core]).mean()
    assert agree > 0.99


def test_band_zero_pins_disparity(oracle):
    """band 0 with a prior constant along each row: a row admits one plane over
    its whole width, so every candidate disparity is the row's rounded prior
    and there is no second candidate. (Where the prior varies along a row the
    row interval is the union over its pixels, so a pixel can see planes its
    own band excludes, as in the reference.)"""
    from doubleeye_amd import synth
    L, R, _ = synth.fronto_pair(200, 100, seed=8)
    rows = np.random.default_rng(8).integers(5, 40, L.shape[0]).astype(np.float32) + 0.3
    prior = np.repeat(rows[:, None], L.shape[1], axis=1)
    d, _, cand, _ = oracle.dense_run_prior(L, R, oracle.cfg(dmax=48), prior, 0, want_cand=True)
    ok = cand[2] > -1e29                       # -1e30: no candidate
    assert ok.mean() > 0.5
    assert np.array_equal(cand[2][ok], np.round(prior)[ok])
    assert not (cand[3] > -1e29).any()         # one admitted plane: no second candidate
