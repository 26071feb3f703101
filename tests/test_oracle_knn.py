"""The keypoint-graph BP oracle (oracle/knn_bp_ref.cpp). The reference has no
such solver, so the oracle is pinned by what the mathematics fixes: min-sum BP
on a tree gives exact min-marginals (checked against enumeration), label
candidates and the k-NN graph against direct enumeration in numpy. CPU only."""
import numpy as np
import pytest

import knn_cases as kc


@pytest.mark.parametrize("seed", range(40))
def test_bp_is_exact_on_trees(seed, oracle):
    from oracle import knn_binding as kb
    rng = np.random.default_rng(seed)
    n = int(rng.integers(1, 8))
    cd, cc, nc, off, adj = kc.random_tree(rng, n)
    c = kb.cfg(lam=float(rng.choice([0.5, 1.0, 2.0])), trunc=int(rng.choice([1, 3, 6, 100])),
               iters=n + 1)
    out, bel = kb.bp_graph(cd, cc, nc, off, adj, c, want_beliefs=True)
    mm = kb.brute_min_marginals(cd, cc, nc, off, adj, c)
    for i in range(n):
        b = bel[i, :nc[i]].astype(np.float64)
        m = mm[i, :nc[i]]
        assert np.array_equal(b - b.min(), m - m.min()), (seed, i, b, m)
        order = sorted(range(nc[i]), key=lambda l: (b[l], l))
        assert out[i]["d1"] == cd[i, order[0]] and out[i]["c1"] == bel[i, order[0]]
        if nc[i] > 1:
            assert out[i]["d2"] == cd[i, order[1]] and out[i]["c2"] == bel[i, order[1]]
        else:
            assert out[i]["d2"] == -1 and np.isinf(out[i]["c2"])


def test_no_smoothness_picks_the_unary_minimum(oracle):
    from oracle import knn_binding as kb
    rng = np.random.default_rng(3)
    cd, cc, nc, off, adj = kc.random_graph(rng, 60)
    out = kb.bp_graph(cd, cc, nc, off, adj, kb.cfg(lam=0.0, iters=5))
    for i in range(60):
        if nc[i] == 0:
            assert out[i]["d1"] == -1
            continue
        l = min(range(nc[i]), key=lambda t: (cc[i, t], t))
        assert out[i]["d1"] == cd[i, l]


def test_candidates_by_enumeration(oracle):
    from doubleeye_amd import synth
    from oracle import knn_binding as kb
    L, R, kp, _ = kc.workload(300, 424, 240)
    c = kb.cfg(dmax=63)
    cd, cc, nc = kb.candidates(L, R, kp, c)
    for i, (x, y) in enumerate(kp):
        a = int(synth.census_at(L, [x], [y], 4, 3)[0])
        ds = [d for d in range(64) if x - d >= 4]
        crs = synth.census_at(R, [x - d for d in ds], [y] * len(ds), 4, 3)
        hd = sorted((bin(a ^ int(b)).count("1"), d) for b, d in zip(crs, ds))[:8]
        assert nc[i] == len(hd)
        assert [d for _, d in hd] == list(cd[i, :nc[i]])
        assert [float(h) for h, _ in hd] == list(cc[i, :nc[i]])


def test_knn_graph_by_enumeration(oracle):
    from oracle import knn_binding as kb
    rng = np.random.default_rng(9)
    kp = np.stack([rng.integers(0, 60, 300), rng.integers(0, 40, 300)], 1).astype(np.int32)
    off, adj = kb.knn_graph(kp, 8)
    d2 = ((kp[:, None, :].astype(np.int64) - kp[None, :, :]) ** 2).sum(-1)
    nb = [set() for _ in range(300)]
    for i in range(300):
        order = sorted((int(d2[i, j]), j) for j in range(300) if j != i)[:8]
        for _, j in order:
            nb[i].add(j)
            nb[j].add(i)
    for i in range(300):
        assert list(adj[off[i]:off[i + 1]]) == sorted(nb[i])


def test_workload_runs(oracle):
    from oracle import knn_binding as kb
    L, R, kp, gt = kc.workload(500, 424, 240)
    out = kb.run(L, R, kp, kb.cfg(dmax=63))
    truth = gt[kp[:, 1], kp[:, 0]]
    assert (out["d1"] == truth).mean() > 0.7   # mostly right on the fronto pair
