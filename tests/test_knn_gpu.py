"""Keypoint-graph BP on the GPU against the oracle (oracle/knn_bp_ref.cpp),
which tests/test_oracle_knn.py pins to exact min-marginals on trees and to
direct enumeration of labels and neighbours. The bar is bit-exact: labels,
graph, beliefs and top-2 outputs, for any lambda (one fixed float order)."""
import sys

import numpy as np
import pytest

import knn_cases as kc

pytestmark = pytest.mark.gpu


def note(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def same_out(a, b):
    return (np.array_equal(a["d1"], b["d1"]) and np.array_equal(a["d2"], b["d2"]) and
            np.array_equal(a["c1"].view(np.uint32), b["c1"].view(np.uint32)) and
            np.array_equal(a["c2"].view(np.uint32), b["c2"].view(np.uint32)))


def gcfg(**kw):
    from doubleeye_amd.knn import KnnConfig
    return KnnConfig(**kw)


def ocfg(**kw):
    from oracle import knn_binding as kb
    return kb.cfg(**kw)


@pytest.mark.parametrize("seed", range(24))
def test_bp_graph_equals_oracle(seed, ctx):
    from doubleeye_amd import knn
    from oracle import knn_binding as kb
    rng = np.random.default_rng(100 + seed)
    if seed % 2:
        cd, cc, nc, off, adj = kc.random_tree(rng, int(rng.integers(1, 30)))
    else:
        cd, cc, nc, off, adj = kc.random_graph(rng, int(rng.integers(2, 300)),
                                               deg=int(rng.integers(1, 9)))
    lam = float(rng.choice([1.0, 0.37, 2.5]))
    tr = int(rng.choice([0, 3, 6, 1000]))
    it = int(rng.integers(0, 12))
    note(f"[knn] graph seed {seed} n={len(nc)} E={len(adj)} lam={lam} trunc={tr} iters={it}")
    g, gb = knn.knn_bp_graph(cd, cc, nc, off, adj, gcfg(lam=lam, trunc=tr, iters=it),
                             want_beliefs=True, ctx=ctx)
    o, ob = kb.bp_graph(cd, cc, nc, off, adj, ocfg(lam=lam, trunc=tr, iters=it),
                        want_beliefs=True)
    assert same_out(g, o)
    assert np.array_equal(gb.view(np.uint32), ob.view(np.uint32))


@pytest.mark.parametrize("shape", [(848, 480, 2000, 127), (424, 240, 700, 63), (200, 60, 33, 100),
                                   (64, 16, 1, 20), (96, 40, 9, 255)])
def test_workload_equals_oracle(shape, ctx):
    from doubleeye_amd import knn
    from oracle import knn_binding as kb
    W, H, n, dmax = shape
    L, R, kp, _ = kc.workload(n, W, H, seed=3, kp_seed=n)
    note(f"[knn] workload {W}x{H} n={n} dmax={dmax}")
    gc, oc = gcfg(dmax=dmax), ocfg(dmax=dmax)
    cd, cc, nc = knn.knn_candidates(L, R, kp, gc, ctx=ctx)
    ocd, occ, onc = kb.candidates(L, R, kp, oc)
    assert np.array_equal(nc, onc) and np.array_equal(cd, ocd)
    assert np.array_equal(cc.view(np.uint32), occ.view(np.uint32))
    off, adj = knn.knn_graph(L, R, kp, gc, ctx=ctx)
    ooff, oadj = kb.knn_graph(kp, 8)
    assert np.array_equal(off, ooff) and np.array_equal(adj, oadj)
    assert same_out(knn.knn_bp(L, R, kp, gc, ctx=ctx), kb.run(L, R, kp, oc))


def test_config_variants_equal_oracle(ctx):
    from doubleeye_amd import knn
    from oracle import knn_binding as kb
    L, R, kp, _ = kc.workload(900, 424, 240, seed=5, kp_seed=9)
    for kw in [dict(census_hw=3), dict(n_cand=3, k_nn=4), dict(k_nn=0), dict(n_cand=1, iters=3),
               dict(w_data=0.25, lam=0.7, trunc=2), dict(iters=0)]:
        note(f"[knn] cfg {kw}")
        assert same_out(knn.knn_bp(L, R, kp, gcfg(**kw), ctx=ctx), kb.run(L, R, kp, ocfg(**kw)))


def test_batch_equals_single(ctx):
    from doubleeye_amd import knn
    shapes = [(848, 480, 2000), (424, 240, 500), (320, 240, 1), (640, 360, 1200), (320, 240, 0)]
    frames, want = [], []
    for k, (W, H, n) in enumerate(shapes):
        L, R, kp, _ = kc.workload(max(n, 1), W, H, seed=7 + k, kp_seed=k)
        kp = kp[:n]
        want.append(knn.knn_bp(L, R, kp, gcfg(), ctx=ctx))
        frames.append(knn.KnnFrame(ctx.to_device(L), ctx.to_device(R),
                                   ctx.to_device(kp) if n else ctx.empty((0, 2), np.int32),
                                   ctx.empty((max(n, 1),), knn.OUT)))
    note("[knn] batch of 5")
    knn.knn_bp_batch_dev(ctx, frames, [gcfg()] * len(frames))
    ctx.sync()
    for k, f in enumerate(frames):
        got = f.out.numpy()[: shapes[k][2]]
        assert same_out(got, want[k]), k


def test_invalid_inputs_raise(ctx):
    from doubleeye_amd import knn
    from doubleeye_amd._lib import DeError
    L, R, kp, _ = kc.workload(10, 100, 50)
    bad = kp.copy()
    bad[3] = (1, 1)    # census window leaves the image
    with pytest.raises(DeError):
        knn.knn_bp(L, R, bad, ctx=ctx)
    with pytest.raises(DeError):
        knn.knn_bp(L, R, kp, gcfg(n_cand=9), ctx=ctx)
    # all keypoints at one position: in-degree beyond the reverse-list capacity
    dup = np.tile(kp[:1], (80, 1))
    with pytest.raises(DeError):
        knn.knn_bp(L, R, dup, ctx=ctx)
