"""The sparse-matcher oracle (oracle/sparse_ref.cpp) against the reference's
own match.cpp: golden outputs written by tests/golden/make_sparse_golden.py
from the compiled reference, exhaustive search, and the assertions of
core/tests/test_match.cpp. CPU only."""
import os

import numpy as np
import pytest

import sparse_cases as sc
from conftest import GOLDEN

CASES = sc.load_golden(os.path.join(GOLDEN, "sparse_match.npz"))


def _cfg(p):
    from oracle import sparse_binding as sb
    return sb.match_cfg(**p["cfg"])


@pytest.mark.parametrize("p", CASES, ids=[p["name"] for p in CASES])
def test_reference_mode_equals_reference(p, oracle):
    from oracle import sparse_binding as sb
    m, st = sb.match_masda(p["L"], p["dl"], p["R"], p["dr"], _cfg(p), canonical=False)
    assert sc.same_matches(m, p["ref_masda"])
    assert sc.stats_tuple(st) == sc.stats_tuple(p["ref_stats"])
    nn, _ = sb.match_mutual_nn(p["L"], p["dl"], p["R"], p["dr"], _cfg(p), p["ratio"],
                               canonical=False)
    assert sc.same_matches(nn, p["ref_nn"])


def test_canonical_equals_reference_without_ties(oracle):
    from oracle import sparse_binding as sb
    checked = 0
    for p in CASES:
        if sc.has_exact_ties(p):
            continue
        m, st = sb.match_masda(p["L"], p["dl"], p["R"], p["dr"], _cfg(p), canonical=True)
        assert sc.same_matches(m, p["ref_masda"]), p["name"]
        assert sc.stats_tuple(st) == sc.stats_tuple(p["ref_stats"]), p["name"]
        checked += 1
    assert checked >= 20


def test_brute_force_matches_reference(oracle):
    from oracle import sparse_binding as sb
    for p in CASES:
        if "ref_brute" not in p:
            continue
        b = sb.match_brute_force(p["L"], p["dl"], p["R"], p["dr"], _cfg(p))
        assert sc.same_matches(b, p["ref_brute"]), p["name"]


def test_reference_assertions_hold_for_oracle(oracle):
    from oracle import sparse_binding as sb
    res = {p["name"]: sb.match_masda(p["L"], p["dl"], p["R"], p["dr"], _cfg(p))[0]
           for p in CASES}
    sc.check_reference_assertions(res, CASES)


def test_candidates_by_brute_enumeration(oracle):
    """The candidate graph against a direct enumeration: every (i, j) in the
    band and disparity window, scored by the closed form, best max_candidates
    by (score desc, y, index)."""
    from oracle import sparse_binding as sb
    for p in CASES[:60]:
        c = p["cfg"]
        ei, ej, es = sb.candidates(p["L"], p["dl"], p["R"], p["dr"], _cfg(p))
        L, R = p["L"].astype(np.float32), p["R"].astype(np.float32)
        K = max(1, c["max_candidates"])
        f32 = np.float32
        exp = []
        for i in range(len(L)):
            band = []
            for j in range(len(R)):
                ylo, yhi = f32(L[i, 1] - f32(c["max_dy"])), f32(L[i, 1] + f32(c["max_dy"]))
                if R[j, 1] < ylo or yhi < R[j, 1]:
                    continue
                d = f32(L[i, 0] - R[j, 0])
                if d < f32(c["min_disparity"]) or d > f32(c["max_disparity"]):
                    continue
                ham = bin(int(p["dl"][i]) ^ int(p["dr"][j])).count("1")
                desc = f32(f32(f32(24.0) - f32(ham)) / f32(24.0))
                sy = f32(c["sigma_y"]) if c["sigma_y"] > 0 else f32(1.0)
                dy = f32(f32(L[i, 1] - R[j, 1]) / sy)
                s = f32(desc - f32(f32(f32(c["w_y"]) * dy) * dy))
                band.append((float(R[j, 1]), j, s))
            band.sort(key=lambda t: (t[0], t[1]))
            if len(band) > K:
                band = sorted(band, key=lambda t: -t[2])[:K]   # stable: ties keep band order
            exp += [(i, j, s) for _, j, s in band]
        assert len(exp) == len(ei), p["name"]
        assert [(int(a), int(b)) for a, b in zip(ei, ej)] == [(a, b) for a, b, _ in exp], p["name"]
        assert (np.array([s for _, _, s in exp], np.float32).view(np.uint32) ==
                es.view(np.uint32)).all(), p["name"]
