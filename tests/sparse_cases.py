"""Sparse matcher problems shared by the golden writer and the tests.

The small hand-built problems restate the situations core/tests/test_match.cpp
exercises (single pair, band/disparity gates, mutual exclusivity, brute-force
agreement on tiny random problems, fully degenerate descriptors, margins,
lambda vs gamma, empty inputs, undamped ties); the inputs are built here from
those descriptions, not taken from the reference's files. Larger seeded
problems come from doubleeye_amd.synth.sparse_problem.
"""
from __future__ import annotations

import numpy as np

# MatchConfig fields in de_match_cfg order, with `lam` for lambda.
CFG_FIELDS = ["max_dy", "min_disparity", "max_disparity", "max_candidates", "sigma_y", "w_y",
              "w_prior", "w_smooth", "iterations", "damping", "lam", "gamma", "converge_eps"]
DEFAULTS = dict(max_dy=2.0, min_disparity=1.0, max_disparity=220.0, max_candidates=24,
                sigma_y=1.0, w_y=1.0, w_prior=1.0, w_smooth=1.0, iterations=20, damping=0.4,
                lam=0.0, gamma=0.0, converge_eps=1e-4)


def desc_with(n_bits: int, offset: int = 0) -> int:
    """A 48-bit descriptor with n_bits consecutive bits set (cyclic from offset)."""
    d = 0
    for b in range(n_bits):
        d |= 1 << ((b + offset) % 48)
    return d


def plain(**kw) -> dict:
    """The isolating configuration of the reference's matcher tests: descriptor
    term only, 60 damped iterations, lambda = gamma = 0, disparities 1..100."""
    c = dict(DEFAULTS, max_dy=2.0, min_disparity=1.0, max_disparity=100.0, w_y=0.0,
             iterations=60, damping=0.4, lam=0.0, gamma=0.0)
    c.update(kw)
    return c


def cfg_array(c: dict) -> np.ndarray:
    return np.array([float(c[k]) for k in CFG_FIELDS], np.float64)


def cfg_from_array(a) -> dict:
    c = {k: float(v) for k, v in zip(CFG_FIELDS, a)}
    c["max_candidates"] = int(c["max_candidates"])
    c["iterations"] = int(c["iterations"])
    return c


def _prob(name, L, dl, R, dr, cfg, ratio=0.85):
    return dict(name=name, L=np.asarray(L, np.float32).reshape(-1, 2),
                dl=np.asarray(dl, np.uint64), R=np.asarray(R, np.float32).reshape(-1, 2),
                dr=np.asarray(dr, np.uint64), cfg=cfg, ratio=ratio)


def small_cases() -> list:
    out = []
    d20 = desc_with(20)
    out.append(_prob("single_pair", [(100, 50)], [d20], [(80, 50)], [d20], plain()))
    out.append(_prob("gate_dy", [(100, 50)], [d20], [(80, 60)], [d20], plain()))
    out.append(_prob("gate_negative_d", [(100, 50)], [d20], [(120, 50)], [d20], plain()))
    out.append(_prob("gate_large_d", [(100, 50)], [d20], [(-50, 50)], [d20], plain()))
    out.append(_prob("gate_y_penalty", [(100, 50)], [d20], [(80, 51.5)], [d20],
                     plain(w_y=1.0, sigma_y=0.5)))
    out.append(_prob("mutual_exclusivity", [(100, 50), (101, 50)], [d20, d20], [(80, 50)],
                     [d20], plain()))
    # fully degenerate: identical descriptors everywhere
    n = 6
    out.append(_prob("degenerate", [(100 + 10 * i, 50) for i in range(n)], [d20] * n,
                     [(80 + 10 * i, 50) for i in range(n)], [d20] * n, plain()))
    # convergence on a well-conditioned chain and undamped on identical descriptors
    n = 12
    Lc = [(120 + 7 * i, 50) for i in range(n)]
    Rc = [(100 + 7 * i, 50) for i in range(n)]
    out.append(_prob("converges", Lc, [desc_with(18, i) for i in range(n)], Rc,
                     [desc_with(18, i) for i in range(n)], plain()))
    out.append(_prob("undamped_ties", Lc, [d20] * n, Rc, [d20] * n,
                     plain(damping=0.0, iterations=40)))
    # margins (lambda = gamma = -0.1)
    d24 = desc_with(24)
    out.append(_prob("margin_clear", [(50, 20)], [d24], [(30, 20), (20, 20)],
                     [d24, desc_with(24, 12)], plain(lam=-0.1, gamma=-0.1)))
    out.append(_prob("margin_tied", [(50, 20)], [d24], [(30, 20), (20, 20)], [d24, d24],
                     plain(lam=-0.1, gamma=-0.1)))
    out.append(_prob("margin_single", [(50, 20)], [d24], [(30, 20)], [d24],
                     plain(lam=-0.1, gamma=-0.1)))
    # lambda vs gamma on a chance-scoring pair (Hamming == 24 of 48 bits)
    chance = desc_with(24, 12)
    out.append(_prob("lambda_charged", [(50, 20), (51, 20)], [d24, d24], [(30, 20)], [chance],
                     plain(lam=-1.0, gamma=0.0)))
    out.append(_prob("gamma_not_charged", [(50, 20), (51, 20)], [d24, d24], [(30, 20)],
                     [chance], plain(lam=0.0, gamma=-1.0)))
    # empty inputs
    one = [(50, 50)]
    out.append(_prob("empty_empty", np.zeros((0, 2)), [], np.zeros((0, 2)), [], plain()))
    out.append(_prob("one_empty", one, [desc_with(10)], np.zeros((0, 2)), [], plain()))
    out.append(_prob("empty_one", np.zeros((0, 2)), [], one, [desc_with(10)], plain()))
    return out


def brute_force_cases(trials: int = 60, seed: int = 12345) -> list:
    """Tiny random problems (3..5 left, 3..5 right keypoints on one row) with
    lambda = 0.05, for agreement with exhaustive maximum-weight matching."""
    rng = np.random.default_rng(seed)
    out = []
    for t in range(trials):
        n = 3 + t % 3
        m = 3 + (t // 3) % 3
        L = [(rng.uniform(20, 120) + 60.0, 50.0) for _ in range(n)]
        dl = [desc_with(int(rng.integers(6, 31)), int(rng.integers(0, 48))) for _ in range(n)]
        R = [(rng.uniform(20, 120), 50.0) for _ in range(m)]
        dr = [desc_with(int(rng.integers(6, 31)), int(rng.integers(0, 48))) for _ in range(m)]
        out.append(_prob(f"brute_{t}", L, dl, R, dr, plain(lam=0.05, gamma=0.0)))
    return out


def random_cases(count: int = 40, seed: int = 2024) -> list:
    """Random problems, half with heavy exact ties (integer positions, few
    distinct descriptors), half tie-free (sub-pixel positions, random 48-bit
    descriptors), with varied configurations."""
    rng = np.random.default_rng(seed)
    out = []
    for t in range(count):
        n = int(rng.integers(1, 120))
        m = int(rng.integers(1, 120))
        ties = t % 2 == 0
        if ties:
            L = np.stack([rng.integers(20, 300, n), rng.integers(0, 8, n)], 1)
            R = np.stack([rng.integers(0, 280, m), rng.integers(0, 8, m)], 1)
            dl = [desc_with(int(rng.integers(6, 30)), int(rng.integers(0, 4))) for _ in range(n)]
            dr = [desc_with(int(rng.integers(6, 30)), int(rng.integers(0, 4))) for _ in range(m)]
        else:
            L = np.stack([rng.uniform(20, 300, n), rng.uniform(0, 8, n)], 1)
            R = np.stack([rng.uniform(0, 280, m), rng.uniform(0, 8, m)], 1)
            dl = rng.integers(0, 2 ** 48, n, dtype=np.uint64)
            dr = rng.integers(0, 2 ** 48, m, dtype=np.uint64)
        cfg = dict(DEFAULTS, max_candidates=int(rng.integers(1, 12)),
                   w_y=float(rng.choice([0.0, 1.0])), lam=float(rng.choice([0.0, 0.05, -0.1])),
                   gamma=float(rng.choice([0.0, -0.1])), iterations=int(rng.integers(0, 40)),
                   damping=float(rng.choice([0.0, 0.4, 0.7])),
                   max_dy=float(rng.choice([1.0, 2.0, 3.5])))
        out.append(_prob(f"random_{t}_{'ties' if ties else 'free'}", L, dl, R, dr, cfg,
                         ratio=float(rng.choice([0.85, 0.7]))))
    return out


def large_cases() -> list:
    """Seeded keypoint problems on the 848x480 random-dot pair."""
    from doubleeye_amd import synth
    out = []
    for seed, n, k, it in [(3, 2000, 8, 10), (5, 1200, 24, 20)]:
        lxy, dl, rxy, dr, _ = synth.sparse_problem(n, 848, 480, seed)
        out.append(_prob(f"synth_{n}_k{k}", lxy, dl, rxy, dr,
                         dict(DEFAULTS, max_candidates=k, iterations=it)))
    return out


def has_exact_ties(p: dict) -> bool:
    """True when an order among exactly equal values may decide the result:
    right keypoints sharing a y, or two band/window pairs of the problem with
    equal scores (the kept-candidate cut, the in-row edge order and the greedy
    completion order all compare scores). Then only the canonical tie-break is
    compared exactly. Uses the oracle's candidate graph with no cap."""
    from oracle import sparse_binding as sb
    if len(np.unique(p["R"][:, 1])) < len(p["R"]):
        return True
    c = dict(p["cfg"], max_candidates=max(1, len(p["R"])))
    _, _, es = sb.candidates(p["L"], p["dl"], p["R"], p["dr"], sb.match_cfg(**c))
    return len(np.unique(es)) < len(es)


def load_golden(path: str) -> list:
    """Problems and the reference's outputs from sparse_match.npz."""
    z = np.load(path, allow_pickle=False)
    out = []
    for k, name in enumerate(z["names"]):
        p = _prob(str(name), z[f"{k}_L"], z[f"{k}_dl"], z[f"{k}_R"], z[f"{k}_dr"],
                  cfg_from_array(z[f"{k}_cfg"]), float(z[f"{k}_ratio"]))
        p["ref_masda"] = z[f"{k}_masda"]
        p["ref_stats"] = z[f"{k}_stats"]
        p["ref_nn"] = z[f"{k}_nn"]
        if f"{k}_brute" in z.files:
            p["ref_brute"] = z[f"{k}_brute"]
        out.append(p)
    return out


def same_matches(a: np.ndarray, b: np.ndarray) -> bool:
    """Bit-exact equality of two Match arrays (floats compared by bits)."""
    if len(a) != len(b):
        return False
    for f in a.dtype.names:
        x, y = np.asarray(a[f]), np.asarray(b[f])
        if x.dtype.kind == "f":
            if not (x.view(np.uint32) == y.view(np.uint32)).all():
                return False
        elif not (x == y).all():
            return False
    return True


def stats_tuple(st) -> tuple:
    names = ["left", "right", "candidates", "iterations_run", "converged", "final_max_delta",
             "oscillating"]
    if isinstance(st, dict):
        return tuple(float(st[n]) for n in names)
    if isinstance(st, (np.ndarray, tuple, list)):
        return tuple(float(v) for v in st)
    return tuple(float(getattr(st, n)) for n in names)


def one_to_one(m: np.ndarray) -> bool:
    return len(np.unique(m["left"])) == len(m) and len(np.unique(m["right"])) == len(m)


def objective(m: np.ndarray, c: dict, n_left: int, n_right: int) -> float:
    t = np.float32(0.0)
    for s in np.asarray(m["score"], np.float32):
        t = np.float32(t + s)
    t = np.float32(t + np.float32(c["lam"]) * np.float32(n_left - len(m)))
    t = np.float32(t + np.float32(c["gamma"]) * np.float32(n_right - len(m)))
    return float(t)


def check_reference_assertions(results: dict, cases: list) -> None:
    """The checks core/tests/test_match.cpp makes of the matcher, applied to a
    matcher's outputs: results maps a problem name to its match_masda output;
    brute-force optima come from the problems ("ref_brute")."""
    by = {p["name"]: p for p in cases}
    r = results
    m = r["single_pair"]
    assert len(m) == 1 and m[0]["left"] == 0 and m[0]["right"] == 0
    assert abs(m[0]["disparity"] - 20.0) < 1e-4 and abs(m[0]["dy"]) < 1e-6
    for g in ("gate_dy", "gate_negative_d", "gate_large_d", "gate_y_penalty"):
        assert len(r[g]) == 0, g
    assert len(r["mutual_exclusivity"]) == 1
    m = r["degenerate"]
    p = by["degenerate"]
    assert one_to_one(m) and len(m) > 0
    assert objective(m, p["cfg"], 6, 6) >= objective(p["ref_brute"], p["cfg"], 6, 6) - 1e-4
    trials = optimal = 0
    sm = sb_ = 0.0
    for name, p in by.items():
        if not name.startswith("brute_"):
            continue
        a = r[name]
        n, mm = len(p["L"]), len(p["R"])
        va, vb = objective(a, p["cfg"], n, mm), objective(p["ref_brute"], p["cfg"], n, mm)
        assert one_to_one(a), name
        trials += 1
        optimal += va >= vb - 1e-4
        sm += va
        sb_ += vb
    assert optimal >= trials * 9 // 10, (optimal, trials)
    assert sm / sb_ > 0.98
    assert r["margin_clear"][0]["margin"] > 0.3
    assert r["margin_tied"][0]["margin"] < 1e-5
    m = r["margin_single"][0]
    assert abs(m["margin"] - (m["score"] - np.float32(-0.1))) < 1e-5
    assert len(r["lambda_charged"]) == 1 and len(r["gamma_not_charged"]) == 0
    for e in ("empty_empty", "one_empty", "empty_one"):
        assert len(r[e]) == 0
