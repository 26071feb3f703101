"""NumPy restatement of the reference's edge-list MASDA (the article code).

TEST INFRASTRUCTURE ONLY: imported by tests/ and the cpu_baseline leg of
bench.py as the checker. The product package never imports this module.

Follows article/masda_stereo.py and article/dense_masda.py:
  seg_max_excluding   masda_stereo.py:417-442
  masda_sparse        masda_stereo.py:453-492
  masda               masda_stereo.py:329-389 (dense score matrix)
  decide              masda_stereo.py:392-414
  dense_problem       dense_masda.py:39-79 (census from the compiled oracle,
                      equal to the rolled-array census on interior pixels)
  census_points       masda_stereo.py:264-283
  hamming             masda_stereo.py:286-294
  build_problem       masda_stereo.py:302-326
`canonical=False` orders the greedy decision with numpy's default argsort,
as the reference does; `canonical=True` breaks exact belief ties by edge
index (a stable sort), the order the HIP path uses.
"""
from __future__ import annotations

import numpy as np

from . import binding


def seg_max_excluding(vals, idx, n):
    """For each element, the max over the other elements of its segment."""
    vals = np.asarray(vals, np.float64)
    idx = np.asarray(idx, np.int64)
    top = np.full(n, -np.inf)
    np.maximum.at(top, idx, vals)
    at_top = vals >= top[idx]
    ties = np.zeros(n, np.int64)
    np.add.at(ties, idx[at_top], 1)
    runner = np.full(n, -np.inf)
    if (~at_top).any():
        np.maximum.at(runner, idx[~at_top], vals[~at_top])
    other_at_top = np.where(ties > 1, top, runner)
    return np.where(at_top, other_at_top[idx], top[idx])


def _greedy(order, ei, ej, se, lam, m, n, finite_stop=None):
    used_i = np.zeros(m, bool)
    used_j = np.zeros(n, bool)
    out = {}
    for e in order:
        if finite_stop is not None and not finite_stop[e]:
            break
        if se[e] <= lam:
            continue
        i, j = int(ei[e]), int(ej[e])
        if used_i[i] or used_j[j]:
            continue
        out[i] = j
        used_i[i] = used_j[j] = True
    return out


def _order(belief, canonical):
    return np.argsort(-belief, kind="stable") if canonical else np.argsort(-belief)


def masda_sparse(ei, ej, se, m, n, lam=-0.1, gam=-0.1, iters=30, damping=0.4, eps=1e-5,
                 canonical=True, return_belief=False):
    """Edge-list MASDA: (assignment dict, per-iteration max change)."""
    ei = np.asarray(ei, np.int64)
    ej = np.asarray(ej, np.int64)
    s = np.asarray(se)
    beta = np.zeros(len(s))
    rho = np.zeros(len(s))
    hist = []
    for _ in range(iters):
        r_new = (1 - damping) * (s - np.maximum(lam, seg_max_excluding(beta, ei, m))) \
            + damping * rho
        b_new = (1 - damping) * (s - np.maximum(gam, seg_max_excluding(r_new, ej, n))) \
            + damping * beta
        change = max(float(np.abs(r_new - rho).max()), float(np.abs(b_new - beta).max()))
        rho, beta = r_new, b_new
        hist.append(change)
        if change < eps:
            break
    belief = beta + rho - s
    out = _greedy(_order(belief, canonical), ei, ej, s, lam, m, n)
    if return_belief:
        return out, np.array(hist), belief
    return out, np.array(hist)


def masda(S, lam=-0.1, gam=-0.1, iters=30, damping=0.4, eps=1e-5, canonical=True):
    """Dense score matrix form (-inf = no candidate), restated on its row-major
    edges: the dense solver's top-2 along rows and columns is the same
    max-excluding-self over each row's and column's finite entries."""
    S = np.asarray(S, np.float32)
    m, n = S.shape
    ei, ej = np.nonzero(np.isfinite(S))
    se = S[ei, ej]
    if len(se) == 0:
        return {}, np.zeros(0)
    beta = np.zeros(len(se))
    rho = np.zeros(len(se))
    hist = []
    for _ in range(iters):
        r_new = (1 - damping) * (se - np.maximum(lam, seg_max_excluding(beta, ei, m))) \
            + damping * rho
        b_new = (1 - damping) * (se - np.maximum(gam, seg_max_excluding(r_new, ej, n))) \
            + damping * beta
        change = max(float(np.abs(r_new - rho).max()), float(np.abs(b_new - beta).max()))
        rho, beta = r_new, b_new
        hist.append(change)
        if change < eps:
            break
    bel = np.full((m, n), -np.inf)
    bel[ei, ej] = beta + rho - se
    return decide(S, bel, lam, canonical), np.array(hist)


def decide(S, belief, lam, canonical=True):
    """Greedy in belief order over a dense matrix, stopping at the first
    non-finite belief."""
    S = np.asarray(S)
    m, n = S.shape
    flat = np.asarray(belief, np.float64).ravel()
    order = _order(flat, canonical)
    ei, ej = np.divmod(np.arange(m * n), n)
    return _greedy(order, ei, ej, S.ravel(), lam, m, n, finite_stop=np.isfinite(flat))


def dense_problem(left, right, dmin, dmax, census_cfg=(3, 3)):
    """Every interior pixel against every candidate d on its row, d-major."""
    H, W = left.shape
    hw, hh = census_cfg
    cl = binding.census(left, hw, hh)
    cr = binding.census(right, hw, hh)
    xs = np.arange(hw, W - hw)
    ys = np.arange(hh, H - hh)
    ei, ej, se = [], [], []
    for d in range(int(dmin), int(dmax) + 1):
        x = xs[xs - d >= hw]
        if len(x) == 0:
            continue
        gy = np.repeat(ys, len(x))
        gx = np.tile(x, len(ys))
        a = cl[gy, gx] ^ cr[gy, gx - d]
        h = np.array([bin(int(v)).count("1") for v in a], np.int64) if len(a) < 64 else \
            np.unpackbits(a.view(np.uint8).reshape(-1, 8), axis=1).sum(1)
        ei.append((gy * W + gx).astype(np.int64))
        ej.append((gy * W + gx - d).astype(np.int64))
        se.append(((24.0 - h) / 24.0).astype(np.float32))
    if not ei:
        return np.zeros(0, np.int64), np.zeros(0, np.int64), np.zeros(0, np.float32), H * W
    return np.concatenate(ei), np.concatenate(ej), np.concatenate(se), H * W


# ---- census / hamming / build_problem (masda_stereo.py:264-326) ----

BITS = 48   # masda_stereo.py:302


def census_points(img, pts, half_w=3, half_h=3):
    """masda_stereo.py:264-283: a bit per neighbour in scan order, set when it
    is darker than the centre, numpy indexing (negative wraps)."""
    out = np.zeros(len(pts), np.uint64)
    f = np.asarray(img).astype(np.int32)
    for n, (x, y) in enumerate(np.asarray(pts).astype(int)):
        c = f[y, x]
        bits, b = 0, 0
        for dy in range(-half_h, half_h + 1):
            for dx in range(-half_w, half_w + 1):
                if dx == 0 and dy == 0:
                    continue
                if f[y + dy, x + dx] < c:
                    bits |= 1 << b
                b += 1
        out[n] = np.uint64(bits)
    return out


_POPCNT = np.array([bin(i).count("1") for i in range(256)], np.int32)


def hamming(a, b):
    """masda_stereo.py:289-294: byte-table popcount of a ^ b, int32."""
    x = np.bitwise_xor(np.asarray(a, np.uint64), np.asarray(b, np.uint64))
    total = np.zeros(x.shape, np.int32)
    for shift in range(0, 64, 8):
        total += _POPCNT[((x >> np.uint64(shift)) & np.uint64(0xFF)).astype(np.uint8)]
    return total


def build_problem(pl, dl, pr, dr, max_dy=1.5, dmin=1.0, dmax=60.0, sigma_y=1.0, w_y=1.0):
    """masda_stereo.py:305-326, row by row, numpy's own type rules."""
    m, n = len(pl), len(pr)
    S = np.full((m, n), -np.inf, np.float32)
    for i in range(m):
        dy = pl[i, 1] - pr[:, 1]
        d = pl[i, 0] - pr[:, 0]
        ok = (np.abs(dy) <= max_dy) & (d >= dmin) & (d <= dmax)
        if not ok.any():
            continue
        h = hamming(np.full(ok.sum(), dl[i], np.uint64), dr[ok])
        desc = (BITS / 2 - h) / (BITS / 2)
        pen = w_y * (dy[ok] / sigma_y) ** 2
        S[i, ok] = desc - pen
    return S
