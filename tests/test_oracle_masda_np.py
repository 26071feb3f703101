"""The NumPy MASDA oracle (oracle/masda_np.py) against vectors written from
the reference's own article/masda_stereo.py (tests/golden/masda_np.npz), an
enumeration of _seg_max_excluding, an enumeration of dense_problem's edges,
and exhaustive optimal assignments. CPU only."""
import itertools

import numpy as np
import pytest

import masda_np_cases as mc
from conftest import GOLDEN  # noqa: F401

SP, MX, DC = mc.load()


def test_seg_max_excluding_by_enumeration(oracle):
    from oracle import masda_np as mn
    rng = np.random.default_rng(0)
    for _ in range(200):
        E, n = int(rng.integers(1, 30)), int(rng.integers(1, 8))
        idx = rng.integers(0, n, E)
        vals = rng.integers(-3, 4, E).astype(np.float64)   # many ties
        got = mn.seg_max_excluding(vals, idx, n)
        for e in range(E):
            others = [vals[f] for f in range(E) if f != e and idx[f] == idx[e]]
            assert got[e] == (max(others) if others else -np.inf)


@pytest.mark.parametrize("k", range(len(SP)))
def test_masda_sparse_equals_reference(k, oracle):
    from oracle import masda_np as mn
    p = SP[k]
    a, hist, bel = mn.masda_sparse(p["ei"], p["ej"], p["se"], p["m"], p["n"], canonical=False,
                                   return_belief=True, **p["kw"])
    assert np.array_equal(hist, p["hist"])
    assert np.array_equal(mc.pairs(a), p["assign"])           # same dict, same order
    c, _ = mn.masda_sparse(p["ei"], p["ej"], p["se"], p["m"], p["n"], canonical=True, **p["kw"])
    if not mc.has_belief_ties(bel):
        assert np.array_equal(mc.pairs(c), p["assign"])


@pytest.mark.parametrize("k", range(len(MX)))
def test_masda_matrix_equals_reference(k, oracle):
    from oracle import masda_np as mn
    p = MX[k]
    a, hist = mn.masda(p["S"], canonical=False, **p["kw"])
    assert np.array_equal(hist, p["hist"])
    assert np.array_equal(mc.pairs(a), p["assign"])


@pytest.mark.parametrize("k", range(len(DC)))
def test_decide_equals_reference(k, oracle):
    from oracle import masda_np as mn
    p = DC[k]
    assert np.array_equal(mc.pairs(mn.decide(p["S"], p["bel"], p["lam"], canonical=False)),
                          p["assign"])
    assert np.array_equal(mc.pairs(mn.decide(p["S"], p["bel"], p["lam"], canonical=True)),
                          p["assign"])   # normal beliefs: no exact ties


def test_dense_problem_by_enumeration(oracle):
    """Edges, their order and scores from the definition, with a census
    computed here pixel by pixel."""
    from doubleeye_amd import synth
    from oracle import masda_np as mn
    L, R, _ = synth.fronto_pair(40, 16, 3)
    ei, ej, se, nodes = mn.dense_problem(L, R, 1, 6)
    H, W = L.shape
    assert nodes == H * W

    def census(img, x, y):
        b, v = 0, 0
        for dy in range(-3, 4):
            for dx in range(-3, 4):
                if dx or dy:
                    v |= int(img[y + dy, x + dx] < img[y, x]) << b
                    b += 1
        return v

    k = 0
    for d in range(1, 7):
        for y in range(3, H - 3):
            for x in range(3 + d, W - 3):
                h = bin(census(L, x, y) ^ census(R, x - d, y)).count("1")
                assert (ei[k], ej[k]) == (y * W + x, y * W + x - d)
                assert se[k] == np.float32((24.0 - h) / 24.0)
                k += 1
    assert k == len(se)


def test_masda_sparse_finds_the_optimum_on_tiny_problems(oracle):
    """Exhaustive maximum-weight matching (lambda per unmatched left, gamma per
    unmatched right): MASDA reaches it on the large majority of problems."""
    from oracle import masda_np as mn
    rng = np.random.default_rng(4)
    hit = total = 0
    for _ in range(60):
        m, n = int(rng.integers(2, 6)), int(rng.integers(2, 6))
        S = np.where(rng.random((m, n)) < 0.7, rng.uniform(-0.5, 1, (m, n)), -np.inf)
        ei, ej = np.nonzero(np.isfinite(S))
        a, _ = mn.masda_sparse(ei, ej, S[ei, ej], m, n, lam=-0.1, gam=-0.1, iters=60)

        def value(asg):
            return sum(S[i, j] for i, j in asg.items()) - 0.1 * (m - len(asg)) - 0.1 * (n - len(asg))

        best = -np.inf
        for perm in itertools.product(*[[-1] + list(range(n))] * m):
            used = [j for j in perm if j >= 0]
            if len(set(used)) < len(used) or any(j >= 0 and not np.isfinite(S[i, j])
                                                 for i, j in enumerate(perm)):
                continue
            best = max(best, value({i: j for i, j in enumerate(perm) if j >= 0}))
        total += 1
        hit += value(a) >= best - 1e-9
    assert hit >= 0.9 * total, (hit, total)


def test_census_points_hamming_closed_forms(oracle):
    """census at points: equal to the compiled oracle's census plane at
    interior points (pinned by the census tests), numpy's wrap-around at
    negative indices, IndexError past the end; hamming: popcount by bin()."""
    from oracle import masda_np as mn
    rng = np.random.default_rng(4)
    img = rng.integers(0, 256, (30, 41), dtype=np.uint8)
    for hw, hh in ((3, 3), (4, 3), (1, 2)):
        plane = oracle.census(img, hw, hh)
        pts = np.stack([rng.integers(hw, 41 - hw, 50), rng.integers(hh, 30 - hh, 50)], 1)
        pts = pts.astype(np.float32) + rng.random((50, 2), dtype=np.float32) * 0.99
        got = mn.census_points(img, pts, hw, hh)
        assert np.array_equal(got, plane[pts[:, 1].astype(int), pts[:, 0].astype(int)])
        # a point at the top-left corner reads the last rows/columns for its
        # negative offsets: the census of the image rolled so it is interior
        rolled = np.roll(img, (hh, hw), axis=(0, 1))
        assert mn.census_points(img, np.array([[0, 0]]), hw, hh)[0] == \
            oracle.census(rolled, hw, hh)[hh, hw]
        with pytest.raises(IndexError):
            mn.census_points(img, np.array([[41 - hw, 5]]), hw, hh)
    a = rng.integers(0, 2 ** 63, 200, dtype=np.uint64) * np.uint64(2) + np.uint64(1)
    b = rng.integers(0, 2 ** 63, 200, dtype=np.uint64)
    h = mn.hamming(a, b)
    assert h.dtype == np.int32
    assert h.tolist() == [bin(int(x) ^ int(y)).count("1") for x, y in zip(a, b)]
    assert mn.hamming(a[:, None], b[None, :5]).shape == (200, 5)


def test_build_problem_closed_forms():
    """build_problem: -inf exactly where the gate |dy| <= max_dy, dmin <= d <=
    dmax fails (enumerated); the 48-bit scale gives +1 / 0 / -1 at Hamming
    0 / 24 / 48 with no row offset; the penalty is w_y (dy/sigma_y)^2 (dy = 1,
    sigma_y = 1: exactly 1 subtracted; w_y = 0: none)."""
    from oracle import masda_np as mn
    rng = np.random.default_rng(9)
    m, n = 40, 55
    pl = np.stack([rng.uniform(0, 100, m), rng.integers(0, 6, m) + rng.choice([0, .5], m)], 1)
    pr = np.stack([rng.uniform(0, 100, n), rng.integers(0, 6, n) + rng.choice([0, .5], n)], 1)
    pl, pr = pl.astype(np.float32), pr.astype(np.float32)
    dl = rng.integers(0, 2 ** 48, m, dtype=np.uint64)
    dr = rng.integers(0, 2 ** 48, n, dtype=np.uint64)
    S = mn.build_problem(pl, dl, pr, dr, max_dy=1.5, dmin=1.0, dmax=60.0)
    for i in range(m):
        for j in range(n):
            dy, d = float(pl[i, 1]) - float(pr[j, 1]), float(pl[i, 0]) - float(pr[j, 0])
            gate = abs(dy) <= 1.5 and 1.0 <= d <= 60.0
            assert np.isfinite(S[i, j]) == gate
    full = np.uint64(2 ** 48 - 1)
    pl1 = np.array([[50, 3]], np.float32)
    for dyv, pen in ((0.0, 0.0), (1.0, 1.0), (-1.0, 1.0)):
        pr1 = np.array([[40, 3 - dyv]] * 3, np.float32)
        dr1 = np.array([0, (1 << 24) - 1, int(full)], np.uint64)
        s = mn.build_problem(pl1, np.array([0], np.uint64), pr1, dr1)
        assert s[0].tolist() == [1.0 - pen, 0.0 - pen, -1.0 - pen]
        s0 = mn.build_problem(pl1, np.array([0], np.uint64), pr1, dr1, w_y=0.0)
        assert s0[0].tolist() == [1.0, 0.0, -1.0]
