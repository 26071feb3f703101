"""Sparse keypoint MASDA on the GPU (de_match_masda, de_match_mutual_nn,
de_match_candidates, de_match_masda_batch_dev) against the oracle and the
reference's own outputs.

The bar: every problem bit-exact against the oracle's canonical tie-break
(matches, stats, candidate graph); against the reference's own outputs
(golden, written from the compiled core/src/match.cpp) bit-exact wherever no
exact tie can decide an order; and the assertions of core/tests/test_match.cpp
hold for the GPU's results."""
import os
import sys

import numpy as np
import pytest

import sparse_cases as sc
from conftest import GOLDEN

pytestmark = pytest.mark.gpu

CASES = sc.load_golden(os.path.join(GOLDEN, "sparse_match.npz"))


def note(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def gcfg(c: dict):
    from doubleeye_amd.sparse import MatchConfig
    return MatchConfig(**c)


def ocfg(c: dict):
    from oracle import sparse_binding as sb
    return sb.match_cfg(**c)


@pytest.fixture(scope="module")
def gpu_results(ctx):
    from doubleeye_amd import sparse
    out = {}
    for p in CASES:
        note(f"[sparse] {p['name']} {len(p['L'])}x{len(p['R'])}")
        st = sparse.MatchStats()
        m = sparse.match_masda(p["L"], p["dl"], p["R"], p["dr"], gcfg(p["cfg"]), st, ctx=ctx)
        nn = sparse.match_mutual_nn(p["L"], p["dl"], p["R"], p["dr"], gcfg(p["cfg"]),
                                    p["ratio"], ctx=ctx)
        out[p["name"]] = (m, st, nn)
    return out


def test_masda_equals_oracle(gpu_results):
    from oracle import sparse_binding as sb
    for p in CASES:
        m, st, _ = gpu_results[p["name"]]
        om, ost = sb.match_masda(p["L"], p["dl"], p["R"], p["dr"], ocfg(p["cfg"]), canonical=True)
        assert sc.same_matches(m, om), p["name"]
        assert sc.stats_tuple(st) == sc.stats_tuple(ost), p["name"]


def test_masda_equals_reference_without_ties(gpu_results):
    checked = 0
    for p in CASES:
        if sc.has_exact_ties(p):
            continue
        m, st, _ = gpu_results[p["name"]]
        assert sc.same_matches(m, p["ref_masda"]), p["name"]
        assert sc.stats_tuple(st) == sc.stats_tuple(p["ref_stats"]), p["name"]
        checked += 1
    assert checked >= 20


def test_mutual_nn_equals_oracle_and_reference(gpu_results):
    from oracle import sparse_binding as sb
    for p in CASES:
        _, _, nn = gpu_results[p["name"]]
        onn, _ = sb.match_mutual_nn(p["L"], p["dl"], p["R"], p["dr"], ocfg(p["cfg"]), p["ratio"],
                                    canonical=True)
        assert sc.same_matches(nn, onn), p["name"]
        if not sc.has_exact_ties(p):
            assert sc.same_matches(nn, p["ref_nn"]), p["name"]


def test_reference_assertions_hold_on_gpu(gpu_results):
    sc.check_reference_assertions({k: v[0] for k, v in gpu_results.items()}, CASES)
    # convergence is reported on a well-conditioned problem
    assert gpu_results["converges"][1].converged
    assert gpu_results["converges"][1].candidates > 0


def test_candidates_equal_oracle(ctx):
    from doubleeye_amd import sparse
    from oracle import sparse_binding as sb
    for p in CASES:
        ei, ej, es = sparse.match_candidates(p["L"], p["dl"], p["R"], p["dr"], gcfg(p["cfg"]),
                                             ctx=ctx)
        oi, oj, os_ = sb.candidates(p["L"], p["dl"], p["R"], p["dr"], ocfg(p["cfg"]))
        assert (ei == oi).all() and (ej == oj).all(), p["name"]
        assert (es.view(np.uint32) == os_.view(np.uint32)).all(), p["name"]


def test_priors_equal_oracle(ctx):
    """Soft disparity prior (grid) and per-keypoint smoothness prior terms."""
    from doubleeye_amd import sparse, synth
    from oracle import sparse_binding as sb
    lxy, dl, rxy, dr, _ = synth.sparse_problem(600, 848, 480, 11)
    rng = np.random.default_rng(5)
    cols, rows = 14, 8
    prior = dict(cell=64, cols=cols, rows=rows,
                 disparity=rng.uniform(5, 40, cols * rows).astype(np.float32),
                 sigma=rng.choice([0.0, 2.0, 6.0], cols * rows).astype(np.float32),
                 valid=(rng.random(cols * rows) < 0.7).astype(np.uint8))
    smooth = dict(disparity=rng.uniform(5, 40, 500).astype(np.float32),
                  sigma=rng.choice([0.0, 1.0, 3.0], 500).astype(np.float32),
                  valid=(rng.random(500) < 0.6).astype(np.uint8))
    c = dict(sc.DEFAULTS, max_candidates=8, iterations=15, w_prior=0.5, w_smooth=0.3)
    for pr, sm in [(prior, None), (None, smooth), (prior, smooth)]:
        m = sparse.match_masda(lxy, dl, rxy, dr, gcfg(c), prior=pr, smooth=sm, ctx=ctx)
        om, _ = sb.match_masda(lxy, dl, rxy, dr, ocfg(c), prior=pr, smooth=sm)
        assert sc.same_matches(m, om)


def test_batch_equals_single(ctx):
    """Several frames in one device batch give each frame's single-call result."""
    from doubleeye_amd import sparse, synth
    probs = [synth.sparse_problem(n, 848, 480, s) for n, s in
             [(2000, 21), (300, 22), (1, 23), (1200, 24), (0, 25), (800, 26)]]
    cfgs = [gcfg(dict(sc.DEFAULTS, max_candidates=k, iterations=it))
            for k, it in [(8, 10), (24, 20), (8, 5), (4, 30), (8, 10), (16, 0)]]
    frames, want = [], []
    for (lxy, dl, rxy, dr, _), c in zip(probs, cfgs):
        L = sparse.keypoints(lxy)
        R = sparse.keypoints(rxy)
        m = max(1, min(len(L), len(R)))
        f = sparse.SparseFrame(
            left=ctx.to_device(L) if len(L) else ctx.empty((0,), sparse.KEYPOINT),
            desc_left=ctx.to_device(dl) if len(dl) else ctx.empty((0,), np.uint64),
            right=ctx.to_device(R) if len(R) else ctx.empty((0,), sparse.KEYPOINT),
            desc_right=ctx.to_device(dr) if len(dr) else ctx.empty((0,), np.uint64),
            out=ctx.empty((m,), sparse.MATCH), n_out=ctx.empty((1,), np.int32),
            stats=ctx.empty((1,), sparse.STATS))
        frames.append(f)
        st = sparse.MatchStats()
        want.append((sparse.match_masda(L, dl, R, dr, c, st, ctx=ctx), st))
    note("[sparse] batch of 6")
    sparse.match_masda_batch_dev(ctx, frames, cfgs)
    ctx.sync()
    for k, f in enumerate(frames):
        n = int(f.n_out.numpy()[0])
        got = f.out.numpy()[:n]
        assert sc.same_matches(got, want[k][0]), k
        assert sc.stats_tuple(f.stats.numpy()[0].tolist()) == sc.stats_tuple(want[k][1]), k


def test_invalid_arguments_raise(ctx):
    from doubleeye_amd import sparse
    from doubleeye_amd._lib import DeError
    L = np.zeros((1, 2), np.float32)
    R = np.zeros((9000, 2), np.float32)
    with pytest.raises(DeError):
        sparse.match_masda(L, [0], R, np.zeros(9000, np.uint64), ctx=ctx)
    with pytest.raises(DeError):
        sparse.match_masda(L, [0], L, [0], gcfg(dict(sc.DEFAULTS, iterations=-1)), ctx=ctx)


@pytest.mark.parametrize("nr", [1024, 1025, 2048, 2049, 2600])
def test_sort_sizes_equal_oracle(ctx, nr):
    """Right-keypoint counts on both sides of the sort's forms: up to 2048 keys
    in registers (1024 and 2048 fill the network exactly, 1025 just passes a
    power of two), past that the LDS network."""
    from doubleeye_amd import sparse, synth
    from oracle import sparse_binding as sb
    lxy, dl, rxy, dr, _ = synth.sparse_problem(nr, 848, 480, 40 + nr, outliers=0.25)
    assert len(rxy) >= nr
    rxy, dr = np.ascontiguousarray(rxy[:nr]), np.ascontiguousarray(dr[:nr])
    note(f"[sparse sort] {len(lxy)}x{len(rxy)}")
    c = dict(sc.DEFAULTS, max_candidates=8, iterations=6)
    st = sparse.MatchStats()
    m = sparse.match_masda(lxy, dl, rxy, dr, gcfg(c), st, ctx=ctx)
    om, ost = sb.match_masda(lxy, dl, rxy, dr, ocfg(c), canonical=True)
    assert sc.same_matches(m, om)
    assert sc.stats_tuple(st) == sc.stats_tuple(ost)
