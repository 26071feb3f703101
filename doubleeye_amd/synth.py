"""Seeded synthetic rectified stereo pairs (uint8), numpy only.

Two scene generators:

* ``fronto_pair`` -- random-dot texture (2 px dots, uniform 0-255, 50% density,
  Gaussian blur sigma 0.8) on fronto-parallel planes at integer disparities;
  the right view is the left scene shifted per plane plus N(0, noise) noise.
* ``slanted_pair`` -- the same texture on slanted planes whose disparity
  varies linearly in x and y, right view resampled bilinearly at sub-pixel
  positions, a fraction of the right image replaced by occluding strips of
  unrelated texture, N(0, noise) noise.

Both views sample the same per-plane surface texture, so occlusion is handled
by depth order (the larger disparity is nearer and wins). The ground-truth
left disparity is returned with the pair.
"""
from __future__ import annotations

import numpy as np

__all__ = ["random_dots", "fronto_pair", "slanted_pair", "pair_for", "census_at",
           "sparse_problem"]


def _gauss_kernel(sigma: float) -> np.ndarray:
    r = max(1, int(np.ceil(3.0 * sigma)))
    t = np.arange(-r, r + 1, dtype=np.float64)
    k = np.exp(-0.5 * (t / sigma) ** 2)
    return k / k.sum()


def _blur(img: np.ndarray, sigma: float) -> np.ndarray:
    k = _gauss_kernel(sigma)
    r = len(k) // 2
    p = np.pad(img, ((r, r), (r, r)), mode="reflect")
    tmp = sum(k[i] * p[:, i:i + img.shape[1]] for i in range(len(k)))
    return sum(k[i] * tmp[i:i + img.shape[0], :] for i in range(len(k)))


def random_dots(rng: np.random.Generator, h: int, w: int, dot: int = 2,
                density: float = 0.5, sigma: float = 0.8) -> np.ndarray:
    """Float texture in [0, 255]: dot x dot cells, each lit with probability
    `density` at a uniform 0-255 level, then Gaussian-blurred."""
    gh, gw = -(-h // dot), -(-w // dot)
    lit = rng.random((gh, gw)) < density
    lvl = rng.integers(0, 256, size=(gh, gw)).astype(np.float64)
    g = np.where(lit, lvl, 0.0)
    img = np.repeat(np.repeat(g, dot, axis=0), dot, axis=1)[:h, :w]
    return _blur(img, sigma)


def _sample_rows(tex: np.ndarray, u: np.ndarray) -> np.ndarray:
    """Linear interpolation of each row of `tex` at fractional columns `u`."""
    w = tex.shape[1]
    u = np.clip(u, 0.0, w - 1.000001)
    i0 = np.floor(u).astype(np.int64)
    f = u - i0
    rows = np.arange(tex.shape[0])[:, None]
    return tex[rows, i0] * (1.0 - f) + tex[rows, i0 + 1] * f


def _render(planes, h: int, w: int, rng: np.random.Generator, margin: int):
    """planes: list of (mask_fn(u, y) -> bool array, a, b, c) with disparity
    d(u, y) = a + b*u + c*y in left coordinates u. Returns (L, R, gt)."""
    ys = np.arange(h, dtype=np.float64)[:, None]
    xs = np.arange(w, dtype=np.float64)[None, :]
    texw = w + margin
    L = np.zeros((h, w))
    R = np.zeros((h, w))
    gt = np.zeros((h, w))
    bestL = np.full((h, w), -np.inf)
    bestR = np.full((h, w), -np.inf)
    for mask_fn, a, b, c in planes:
        tex = random_dots(rng, h, texw)
        # left view: surface point u = x
        dL = a + b * xs + c * ys
        vis = mask_fn(xs + 0 * ys, ys + 0 * xs) & (dL > bestL)
        L = np.where(vis, _sample_rows(tex, (xs + 0 * ys)), L)
        gt = np.where(vis, dL, gt)
        bestL = np.where(vis, dL, bestL)
        # right view: solve u - d(u, y) = xr for u
        u = (xs + a + c * ys) / (1.0 - b)
        dR = a + b * u + c * ys
        visR = mask_fn(u, ys + 0 * xs) & (dR > bestR) & (u < texw - 1)
        R = np.where(visR, _sample_rows(tex, u), R)
        bestR = np.where(visR, dR, bestR)
    # right pixels no plane reaches (disoccluded at the left border): texture
    hole = ~np.isfinite(bestR)
    if hole.any():
        R = np.where(hole, random_dots(rng, h, w), R)
    return L, R, gt


def _to_u8(img: np.ndarray) -> np.ndarray:
    return np.clip(np.rint(img), 0, 255).astype(np.uint8)


def fronto_pair(w: int = 424, h: int = 240, seed: int = 1,
                disparities=(8, 20, 35), noise: float = 2.0):
    """Fronto-parallel planes: the farthest covers the frame, nearer ones are a
    rectangle and an ellipse. Returns (left u8, right u8, gt float32)."""
    rng = np.random.default_rng(seed)
    d = sorted(disparities)
    planes = [(lambda u, y: np.ones(np.broadcast(u, y).shape, bool), float(d[0]), 0.0, 0.0)]
    if len(d) > 1:
        x0, x1 = 0.18 * w, 0.55 * w
        y0, y1 = 0.15 * h, 0.70 * h
        planes.append((lambda u, y, x0=x0, x1=x1, y0=y0, y1=y1:
                       (u >= x0) & (u < x1) & (y >= y0) & (y < y1), float(d[1]), 0.0, 0.0))
    if len(d) > 2:
        cx, cy, rx, ry = 0.66 * w, 0.55 * h, 0.16 * w, 0.28 * h
        planes.append((lambda u, y, cx=cx, cy=cy, rx=rx, ry=ry:
                       ((u - cx) / rx) ** 2 + ((y - cy) / ry) ** 2 < 1.0, float(d[2]), 0.0, 0.0))
    L, R, gt = _render(planes, h, w, rng, margin=int(d[-1]) + 2)
    R = R + rng.normal(0.0, noise, size=R.shape)
    return _to_u8(L), _to_u8(R), gt.astype(np.float32)


def slanted_pair(w: int = 848, h: int = 480, seed: int = 2, dlo: float = 4.0,
                 dhi: float = 110.0, nplanes: int = 6, occluded: float = 0.05,
                 noise: float = 3.0):
    """Slanted planes with disparity linear in (x, y) inside [dlo, dhi],
    sub-pixel right view, `occluded` fraction of right columns replaced by
    unrelated texture strips. Returns (left u8, right u8, gt float32)."""
    rng = np.random.default_rng(seed)
    span = dhi - dlo
    planes = []
    # background: far, gently slanted across the frame
    a0 = dlo
    bx = (0.20 * span) / w
    cy = (0.05 * span) / h
    planes.append((lambda u, y: np.ones(np.broadcast(u, y).shape, bool), a0, bx, cy))
    for p in range(1, nplanes):
        # nearer planes get higher disparity bands, each spanning ~15% of the range
        base = dlo + span * (0.15 + 0.70 * p / max(1, nplanes - 1)) - 0.15 * span
        base = min(max(base, dlo), dhi - 0.15 * span)
        cxr = rng.uniform(0.15, 0.85) * w
        cyr = rng.uniform(0.2, 0.8) * h
        rx = rng.uniform(0.08, 0.18) * w
        ry = rng.uniform(0.12, 0.28) * h
        sx = rng.uniform(-0.5, 0.5) * 0.15 * span / (2 * rx)
        sy = rng.uniform(-0.5, 0.5) * 0.15 * span / (2 * ry)
        # d = base + 0.075 span at the centre, +-0.075 span across the ellipse
        a = base + 0.075 * span - sx * cxr - sy * cyr
        planes.append((lambda u, y, cx=cxr, cy_=cyr, rx=rx, ry=ry:
                       ((u - cx) / rx) ** 2 + ((y - cy_) / ry) ** 2 < 1.0, a, sx, sy))
    L, R, gt = _render(planes, h, w, rng, margin=int(dhi) + 4)
    if occluded > 0:
        nstrip = max(1, int(round(occluded * w / 12)))
        widths = np.full(nstrip, occluded * w / nstrip)
        starts = rng.uniform(0, w - widths.max(), size=nstrip)
        occ = random_dots(rng, h, w)
        cols = np.zeros(w, bool)
        for s, wd in zip(starts, widths):
            cols[int(s):int(s + wd)] = True
        R = np.where(cols[None, :], occ, R)
    R = R + rng.normal(0.0, noise, size=R.shape)
    return _to_u8(L), _to_u8(R), gt.astype(np.float32)


def pair_for(w: int, h: int, dmax: int, seed: int, kind: str = "fronto"):
    """A pair whose disparities fit a search range of `dmax`: fronto planes at
    8/20/35 scaled by dmax/48, or slanted planes over [4, 0.86*dmax]."""
    if kind == "fronto":
        s = dmax / 48.0
        ds = tuple(max(1, int(round(v * s))) for v in (8, 20, 35))
        return fronto_pair(w, h, seed, ds)
    return slanted_pair(w, h, seed, 4.0, 0.86 * dmax)


def census_at(img: np.ndarray, xs: np.ndarray, ys: np.ndarray, hw: int = 3,
              hh: int = 3) -> np.ndarray:
    """Census descriptors at integer pixel positions (numpy): bit b is set when
    the b-th neighbour, dy-major then dx-minor with the centre skipped, is darker
    than the centre (the definition of core/src/preproc.cpp:110-162). Points
    whose window crosses the border get 0."""
    img = np.asarray(img, np.int32)
    H, W = img.shape
    xs = np.asarray(xs, np.int64)
    ys = np.asarray(ys, np.int64)
    inside = (xs >= hw) & (xs < W - hw) & (ys >= hh) & (ys < H - hh)
    xc = np.clip(xs, hw, W - hw - 1)
    yc = np.clip(ys, hh, H - hh - 1)
    c = img[yc, xc]
    d = np.zeros(len(xs), np.uint64)
    b = 0
    for dy in range(-hh, hh + 1):
        for dx in range(-hw, hw + 1):
            if dx == 0 and dy == 0:
                continue
            bit = (img[yc + dy, xc + dx] < c).astype(np.uint64) << np.uint64(b)
            d |= bit
            b += 1
    d[~inside] = 0
    return d


def sparse_problem(n: int = 2000, w: int = 848, h: int = 480, seed: int = 3,
                   keep: float = 0.9, outliers: float = 0.1, jitter: float = 0.3):
    """A keypoint matching problem on the fronto random-dot pair: n left
    keypoints drawn uniformly (sub-pixel positions) with 7x7 census
    descriptors; right keypoints are their true correspondences (disparity from
    the ground truth) with N(0, jitter) position noise, a `keep` fraction of
    them, plus `outliers` * n unrelated right keypoints, in shuffled order.

    Returns (left_xy, desc_left, right_xy, desc_right, truth) where truth[k] is
    the left index of right keypoint k or -1."""
    L, R, gt = fronto_pair(w, h, seed)
    rng = np.random.default_rng(seed + 7919)
    border = 8
    lx = rng.uniform(border, w - border, n).astype(np.float32)
    ly = rng.uniform(border, h - border, n).astype(np.float32)
    dl = census_at(L, np.rint(lx), np.rint(ly))
    d = gt[np.rint(ly).astype(int), np.rint(lx).astype(int)].astype(np.float32)
    sel = np.nonzero(rng.random(n) < keep)[0]
    rx = lx[sel] - d[sel] + rng.normal(0, jitter, len(sel)).astype(np.float32)
    ry = ly[sel] + rng.normal(0, jitter, len(sel)).astype(np.float32)
    no = int(round(outliers * n))
    ox = rng.uniform(border, w - border, no).astype(np.float32)
    oy = rng.uniform(border, h - border, no).astype(np.float32)
    rx = np.concatenate([rx, ox]).astype(np.float32)
    ry = np.concatenate([ry, oy]).astype(np.float32)
    truth = np.concatenate([sel, -np.ones(no, np.int64)])
    ok = (rx >= border) & (rx < w - border) & (ry >= border) & (ry < h - border)
    rx, ry, truth = rx[ok], ry[ok], truth[ok]
    perm = rng.permutation(len(rx))
    rx, ry, truth = rx[perm], ry[perm], truth[perm]
    dr = census_at(R, np.rint(rx), np.rint(ry))
    return (np.stack([lx, ly], 1), dl, np.stack([rx, ry], 1), dr, truth)
