"""The keypoint-detector oracle (oracle/detect_ref.cpp) against the reference:
golden outputs of its detect_keypoints_fast and of its de_pipe tool on the six
IR scenes (tests/golden/make_detect_golden.py), the assertions of
core/tests/test_preproc.cpp:176-243, and independent statements of each stage
(FAST by enumerating arcs, Shi-Tomasi by the closed-form eigenvalue in float64,
the texture floor by numpy's std, the suppression by a brute-force greedy, the
refinement's parabola vertex). CPU only."""
import os

import numpy as np
import pytest

import detect_cases as dc
from conftest import ROOT

REF_BUILT = os.path.exists(os.path.join(ROOT, "oracle", "_ref", "libref_sparse.so"))
needs_ref = pytest.mark.skipif(not REF_BUILT, reason="reference harness not built here")


def _db():
    from oracle import detect_binding as db
    return db


def uniform(w, h, v):
    return np.full((h, w), v, np.uint8)


# --- golden vectors from the reference ----------------------------------------

@pytest.mark.parametrize("img,cfg", dc.det_cases())
def test_reference_mode_equals_golden(img, cfg, oracle):
    db = _db()
    want, counts, _ = dc.det_golden(img, cfg)
    got, nc, ns = db.detect_keypoints_fast(dc.image(img), db.detector_cfg(**dc.DET_CFGS[cfg]),
                                           canonical=False, counts=True)
    assert dc.same_bits(dc.as_rows(got), want)
    assert (nc, ns) == counts


def test_canonical_equals_golden_where_ties_do_not_decide(oracle):
    db = _db()
    checked = 0
    for img, cfg in dc.det_cases():
        want, _, canon = dc.det_golden(img, cfg)
        got = db.detect_keypoints_fast(dc.image(img), db.detector_cfg(**dc.DET_CFGS[cfg]))
        assert dc.same_bits(dc.as_rows(got), want) == canon, (img, cfg)
        checked += canon
    assert checked >= 50


@pytest.mark.parametrize("flags", list(dc.PIPE_FLAGS))
def test_pipeline_equals_de_pipe_records(flags, oracle):
    """detect x2 -> census_at -> match_masda -> margin filter -> refine equals
    the DEMR records the reference's de_pipe wrote for the six scene pairs."""
    db = _db()
    det, dr, mc, kw = db.pipe_args(dc.PIPE_FLAGS[flags])
    sc = dc.scenes()
    nonempty = 0
    for s in dc.SCENES:
        _, _, m, rec = db.pipeline(sc[f"{s}_L"], sc[f"{s}_R"], db.detector_cfg(**det),
                                   db.detector_cfg(**dr), (3, 3), db.pipe_match_cfg(**mc), **kw)
        assert dc.same_bits(rec, dc.golden()[f"pipe_{s}_{flags}"]), s
        nonempty += len(rec) > 0
    assert nonempty >= 5


# --- core/tests/test_preproc.cpp:176-243 -------------------------------------

def test_fast_assertions_of_reference_tests(oracle):
    db = _db()
    flat = uniform(64, 64, 100)
    assert not db.is_fast_corner(flat, 32, 32, 20)
    assert not db.is_fast_corner(flat, 1, 1, 20)
    sq = uniform(64, 64, 30)
    sq[20:44, 20:44] = 220
    assert db.is_fast_corner(sq, 20, 20, 20)
    assert not db.is_fast_corner(sq, 32, 20, 20)
    assert not db.is_fast_corner(sq, 32, 32, 20)
    faint = uniform(64, 64, 100)
    faint[20:44, 20:44] = 110
    assert db.is_fast_corner(faint, 20, 20, 5)
    assert not db.is_fast_corner(faint, 20, 20, 30)


@needs_ref
def test_shi_tomasi_at_agrees_with_dense_response(oracle):
    """test_preproc.cpp:199-209: the sparse response within 1e-2 of the dense
    shi_tomasi_response over the interior; the reference's own at-every-pixel
    values equal the oracle's bit for bit."""
    db = _db()
    sq = uniform(64, 64, 30)
    sq[20:44, 20:44] = 220
    at, dense = db.ref_shi_tomasi_maps(sq, 3)
    ours = np.array([[db.shi_tomasi_at(sq, x, y, 3) for x in range(64)] for y in range(64)],
                    np.float32)
    assert dc.same_bits(ours, at)
    assert np.abs(dense[6:58, 6:58] - ours[6:58, 6:58]).max() <= 1e-2


def grid_of_squares():
    img = uniform(256, 128, 20)
    for cy in range(16, 128, 32):
        for cx in range(16, 256, 32):
            img[cy - 5:cy + 6, cx - 5:cx + 6] = 200
    return img


def test_fast_vs_dense_agreement_assertions(oracle):
    """test_preproc.cpp:212-243 on the oracle: keypoints found, inside the
    image, per-cell cap held; coverage within 20% of the dense detector's
    (dense coverage from the reference build when present)."""
    db = _db()
    img = grid_of_squares()
    cfg = db.detector_cfg(cell=32, per_cell=2, min_local_std=0.5)
    for canonical in (False, True):
        fast = db.detect_keypoints_fast(img, cfg, canonical)
        assert len(fast) > 0
        assert ((fast["x"] >= 0) & (fast["x"] <= 256) & (fast["y"] >= 0) &
                (fast["y"] <= 128)).all()
        assert len(fast) <= 8 * 4 * 2
        if REF_BUILT:
            dense = db.ref_detect_keypoints_dense(img, cfg)
            assert db.cell_occupancy(fast, 256, 128, 32) >= \
                0.8 * db.cell_occupancy(dense, 256, 128, 32)


# --- independent statements of each stage -------------------------------------

RING = [(0, -3), (1, -3), (2, -2), (3, -1), (3, 0), (3, 1), (2, 2), (1, 3), (0, 3), (-1, 3),
        (-2, 2), (-3, 1), (-3, 0), (-3, -1), (-2, -2), (-1, -3)]


def fast_by_arcs(img, thr):
    """FAST-9 by enumerating every arc of 9 consecutive ring pixels (the
    compass pre-test is implied: an arc of 9 holds at least two compass points)."""
    img = img.astype(np.int32)
    H, W = img.shape
    out = np.zeros((H, W), np.uint8)
    for y in range(3, H - 3):
        for x in range(3, W - 3):
            c = img[y, x]
            ring = [img[y + dy, x + dx] for dx, dy in RING]
            bright = [v > c + thr for v in ring]
            dark = [(v < c - thr) and not b for v, b in zip(ring, bright)]
            for mask in (bright, dark):
                if any(all(mask[(s + i) % 16] for i in range(9)) for s in range(16)):
                    out[y, x] = 1
                    break
    return out


def test_fast_matches_arc_enumeration(oracle):
    db = _db()
    rng = np.random.default_rng(5)
    for thr in (4, 12, 40):
        img = (rng.integers(0, 4, (24, 28)) * 60).astype(np.uint8)
        assert np.array_equal(db.fast_map(img, thr), fast_by_arcs(img, thr)), thr


def shi_closed_form(img, x, y, gw):
    img = img.astype(np.float64)
    r = max(1, gw // 2)
    sxx = syy = sxy = 0.0
    for yy in range(y - r, y + r + 1):
        for xx in range(x - r, x + r + 1):
            gx = (img[yy - 1, xx + 1] - img[yy - 1, xx - 1] + 2 * (img[yy, xx + 1] - img[yy, xx - 1])
                  + img[yy + 1, xx + 1] - img[yy + 1, xx - 1]) / 8
            gy = (img[yy + 1, xx - 1] - img[yy - 1, xx - 1] + 2 * (img[yy + 1, xx] - img[yy - 1, xx])
                  + img[yy + 1, xx + 1] - img[yy - 1, xx + 1]) / 8
            sxx += gx * gx
            syy += gy * gy
            sxy += gx * gy
    n = (2 * r + 1) ** 2
    return float(np.linalg.eigvalsh(np.array([[sxx, sxy], [sxy, syy]]) / n)[0])


def test_shi_tomasi_is_the_smaller_eigenvalue(oracle):
    db = _db()
    rng = np.random.default_rng(6)
    img = rng.integers(0, 256, (30, 30)).astype(np.uint8)
    for gw in (3, 5, 7):
        for _ in range(40):
            x, y = (int(v) for v in rng.integers(5, 25, 2))
            want = shi_closed_form(img, x, y, gw)
            got = db.shi_tomasi_at(img, x, y, gw)
            assert abs(got - want) <= 1e-4 * max(1.0, abs(want)), (gw, x, y)
    assert db.shi_tomasi_at(img, 1, 15, 3) == 0.0     # window leaves the image


def test_local_std_is_numpy_std(oracle):
    db = _db()
    rng = np.random.default_rng(7)
    img = rng.integers(0, 256, (20, 20)).astype(np.uint8)
    for x, y in [(0, 0), (10, 10), (19, 5), (3, 19)]:
        win = img[max(0, y - 3):y + 4, max(0, x - 3):x + 4].astype(np.float64)
        assert abs(db.local_std_at(img, x, y) - win.std()) <= 1e-5 * max(1.0, win.std())
    assert db.local_std_at(uniform(9, 9, 50), 4, 4) == 0.0


def greedy_by_brute_force(cands, rad):
    """cands: (key, x, y) sorted strongest first -> kept (x, y) in that order."""
    kept = []
    for _, x, y in cands:
        if all(max(abs(x - kx), abs(y - ky)) > rad for kx, ky in kept):
            kept.append((x, y))
    return kept


def test_detector_stages_by_brute_force(oracle):
    """Candidates = FAST corners inside the border with response above the
    floor; suppression = greedy over (response desc, scan order); the output =
    per cell, the survivors above the texture floor, top per_cell by the same
    order, with the sub-pixel vertex within half a pixel."""
    db = _db()
    rng = np.random.default_rng(8)
    img = np.clip(rng.normal(128, 50, (48, 64)), 0, 255).astype(np.uint8)
    for kw in (dict(fast_threshold=10, cell=16, per_cell=2, nms_radius=2),
               dict(fast_threshold=20, cell=8, per_cell=5, nms_radius=1, subpixel=0)):
        cfg = db.detector_cfg(**kw)
        H, W = img.shape
        b = max(cfg.border, 5)
        fm = fast_by_arcs(img, cfg.fast_threshold)
        cands = []
        for y in range(b, H - b):
            for x in range(b, W - b):
                if fm[y, x]:
                    r = np.float32(db.shi_tomasi_at(img, x, y, cfg.grad_window))
                    if not r < np.float32(cfg.min_response):
                        cands.append((r, x, y))
        cands.sort(key=lambda c: (-c[0], c[2], c[1]))
        kept = greedy_by_brute_force(cands, max(1, cfg.nms_radius))
        resp = {(x, y): r for r, x, y in cands}
        cells = {}
        for x, y in kept:
            if db.local_std_at(img, x, y) >= cfg.min_local_std:
                cells.setdefault((y // cfg.cell, x // cfg.cell), []).append((x, y))
        want = []
        for key in sorted(cells):
            want += cells[key][: cfg.per_cell]
        got = db.detect_keypoints_fast(img, cfg)
        assert len(got) == len(want) and len(want) > 3
        for k, (x, y) in zip(got, want):
            assert abs(k["x"] - x) <= 0.5 and abs(k["y"] - y) <= 0.5
            assert k["response"] == resp[(x, y)]
            if not cfg.subpixel:
                assert k["x"] == x and k["y"] == y


def test_refine_is_the_clamped_parabola_vertex(oracle):
    """Disparity = round(d) + vertex of the parabola through the SAD costs at
    round(d) - 1, round(d), round(d) + 1 (7x7 windows), clamped to +-0.5; left
    unchanged near a border or without a minimum."""
    from oracle import sparse_binding as sb
    db = _db()
    rng = np.random.default_rng(9)
    L = rng.integers(0, 256, (40, 80)).astype(np.uint8)
    R = np.roll(L, -6, axis=1)
    R = np.clip(R.astype(np.int32) + rng.integers(-20, 20, R.shape), 0, 255).astype(np.uint8)
    kl = np.zeros(6, sb.KP_DTYPE)
    kl["x"] = [30.2, 40.6, 50.0, 2.0, 70.4, 45.0]
    kl["y"] = [20.0, 10.4, 30.6, 20.0, 20.0, 38.0]
    m = np.zeros(6, sb.MATCH_DTYPE)
    m["left"] = np.arange(6)
    m["right"] = 0
    m["disparity"] = [6.3, 5.6, 7.4, 6.0, 6.0, 6.0]
    got = db.refine_disparity(L, R, kl, kl[:1], m)

    def sad(xl, y, d):
        return int(np.abs(L[y - 3:y + 4, xl - 3:xl + 4].astype(np.int32)
                          - R[y - 3:y + 4, xl - d - 3:xl - d + 4].astype(np.int32)).sum())
    for k in range(6):
        xl, y = int(kl["x"][k] + 0.5), int(kl["y"][k] + 0.5)
        d0 = int(np.floor(m["disparity"][k] + 0.5))
        inside = xl - 3 >= 0 and xl + 3 < 80 and xl - d0 - 4 >= 0 and y - 3 >= 0 and y + 3 < 40
        if not inside:
            assert got["disparity"][k] == m["disparity"][k]
            continue
        cm, c0, cp = sad(xl, y, d0 - 1), sad(xl, y, d0), sad(xl, y, d0 + 1)
        den = cm - 2 * c0 + cp
        if den <= 0:
            assert got["disparity"][k] == m["disparity"][k]
            continue
        want = np.float32(d0 + min(0.5, max(-0.5, 0.5 * (cm - cp) / den)))
        assert got["disparity"][k] == want
    assert got["disparity"][3] == m["disparity"][3]     # x = 2: window leaves the image


# --- the reference itself ------------------------------------------------------

@needs_ref
def test_oracle_equals_reference_on_random_images(oracle):
    db = _db()
    rng = np.random.default_rng(10)
    for t in range(12):
        H, W = int(rng.integers(12, 90)), int(rng.integers(12, 120))
        img = rng.integers(0, 256, (H, W)).astype(np.uint8)
        if t % 3 == 0:
            img = (img // 64 * 64).astype(np.uint8)     # plateaus: equal responses
        cfg = db.detector_cfg(fast_threshold=int(rng.integers(2, 30)),
                              nms_radius=int(rng.integers(0, 5)),
                              cell=int(rng.integers(2, 40)), per_cell=int(rng.integers(0, 6)),
                              grad_window=int(rng.choice([1, 3, 5, 9])),
                              border=int(rng.integers(0, 12)),
                              min_response=float(rng.choice([-1.0, 0.0, 1.0, 30.0])),
                              min_local_std=float(rng.choice([0.0, 2.0, 20.0])),
                              subpixel=int(rng.integers(0, 2)))
        a, nc, ns = db.detect_keypoints_fast(img, cfg, canonical=False, counts=True)
        b, rc, rs = db.ref_detect_keypoints_fast(img, cfg, counts=True)
        assert dc.same_bits(dc.as_rows(a), dc.as_rows(b)), t
        assert (nc, ns) == (rc, rs)


@needs_ref
def test_point_functions_equal_reference(oracle):
    db = _db()
    rng = np.random.default_rng(11)
    img = rng.integers(0, 256, (20, 24)).astype(np.uint8)
    for gw in (3, 5):
        at, _ = db.ref_shi_tomasi_maps(img, gw)
        ours = np.array([[db.shi_tomasi_at(img, x, y, gw) for x in range(24)] for y in range(20)],
                        np.float32)
        assert dc.same_bits(ours, at)
    for thr in (0, 7, 25):
        ref = np.array([[db.ref_is_fast_corner(img, x, y, thr) for x in range(24)]
                        for y in range(20)], np.uint8)
        assert np.array_equal(db.fast_map(img, thr), ref)


@needs_ref
def test_refine_equals_reference(oracle):
    from oracle import sparse_binding as sb
    db = _db()
    rng = np.random.default_rng(12)
    L = rng.integers(0, 256, (60, 90)).astype(np.uint8)
    R = np.roll(L, -9, axis=1)
    n = 200
    kl = np.zeros(n, sb.KP_DTYPE)
    kl["x"] = rng.uniform(-1, 91, n)
    kl["y"] = rng.uniform(-1, 61, n)
    m = np.zeros(n, sb.MATCH_DTYPE)
    m["left"] = rng.integers(-1, n + 1, n)
    m["right"] = rng.integers(-1, 3, n)
    m["disparity"] = rng.uniform(-3, 20, n)
    for half in (1, 3):
        a = db.refine_disparity(L, R, kl, kl[:2], m, half)
        b = db.ref_refine_disparity(L, R, kl, kl[:2], m, half)
        assert dc.same_bits(a["disparity"], b["disparity"])
