"""GPU keypoint detector and detect -> match pipeline (de_detect_keypoints_fast,
de_detect_keypoints_fast_batch_dev, de_fast_score_map, de_census_at_keypoints,
de_refine_disparity, de_sparse_pipeline) against the oracle and the reference's
golden outputs (its detect_keypoints_fast and de_pipe on the six IR scenes).
Bit-exact: positions, responses, texture values, counts, DEMR records."""
import sys

import numpy as np
import pytest

import detect_cases as dc

pytestmark = pytest.mark.gpu


def note(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def gcfg(kw):
    from doubleeye_amd.detect import DetectorConfig
    c = DetectorConfig()
    for k, v in kw.items():
        setattr(c, k, v)
    return c


def ocfg(kw):
    from oracle import detect_binding as db
    return db.detector_cfg(**kw)


def test_fast_and_response_maps_equal_oracle(ctx):
    """is_fast_corner and shi_tomasi_at at every pixel, grad windows 1..7
    (radius 3 takes the global-memory path)."""
    from doubleeye_amd import detect
    from oracle import detect_binding as db
    rng = np.random.default_rng(31)
    imgs = [dc.image("blobs")[:40, :60], rng.integers(0, 256, (33, 47)).astype(np.uint8),
            dc.image("square")]
    for k, img in enumerate(imgs):
        H, W = img.shape
        for thr, gw in [(8, 3), (20, 1), (4, 5), (12, 7)]:
            note(f"[detect] map image {k} {W}x{H} thr {thr} gw {gw}")
            corner, resp = detect.fast_score_map(img, thr, gw, ctx=ctx)
            assert np.array_equal(corner, db.fast_map(img, thr))
            want = np.array([[db.shi_tomasi_at(img, x, y, gw) for x in range(W)]
                             for y in range(H)], np.float32)
            assert dc.same_bits(resp, want)


@pytest.mark.parametrize("img,cfg", dc.det_cases())
def test_detector_equals_reference_golden(img, cfg, ctx):
    """Every golden case: equal to the oracle's tie order always, and to the
    reference's own output wherever equal responses do not decide (all cases
    but the grid of identical squares and one low-threshold scene)."""
    from doubleeye_amd import detect
    from oracle import detect_binding as db
    note(f"[detect] golden {img} {cfg}")
    want, counts, canon = dc.det_golden(img, cfg)
    prof = detect.DetectProfile()
    got = detect.detect_keypoints_fast(dc.image(img), gcfg(dc.DET_CFGS[cfg]), prof, ctx=ctx)
    assert (prof.candidates, prof.suppressed) == counts
    orc = db.detect_keypoints_fast(dc.image(img), ocfg(dc.DET_CFGS[cfg]))
    assert dc.same_bits(dc.as_rows(got), dc.as_rows(orc))
    if canon:
        assert dc.same_bits(dc.as_rows(got), want)


EDGE = [
    ("noise_r12", lambda r: r.integers(0, 256, (90, 130)), dict(nms_radius=12, per_cell=50)),
    ("noise_t0", lambda r: r.integers(0, 256, (70, 100)),
     dict(fast_threshold=0, min_response=-1.0, min_local_std=0.0, border=0)),
    ("checker", lambda r: (np.indices((64, 80)).sum(0) % 2) * 200 + 20,
     dict(fast_threshold=5, min_local_std=0.0, cell=4, per_cell=1)),
    ("plateaus", lambda r: r.integers(0, 4, (80, 96)) * 80, dict(fast_threshold=10, nms_radius=1)),
    ("flat", lambda r: np.full((40, 50), 77), {}),
    ("tiny", lambda r: r.integers(0, 256, (9, 11)), dict(border=0)),
    ("one_px", lambda r: np.full((1, 1), 5), {}),
    ("gw21", lambda r: r.integers(0, 256, (64, 64)), dict(grad_window=21, border=2)),
    ("wide_cells", lambda r: r.integers(0, 256, (100, 300)), dict(cell=200, per_cell=1000)),
    ("neg_thr", lambda r: r.integers(0, 256, (40, 40)), dict(fast_threshold=-3, per_cell=7)),
]


@pytest.mark.parametrize("name,make,kw", EDGE, ids=[e[0] for e in EDGE])
def test_detector_edge_cases_equal_oracle(name, make, kw, ctx):
    from doubleeye_amd import detect
    from oracle import detect_binding as db
    img = make(np.random.default_rng(32)).astype(np.uint8)
    note(f"[detect] edge {name} {img.shape}")
    prof = detect.DetectProfile()
    got = detect.detect_keypoints_fast(img, gcfg(kw), prof, ctx=ctx)
    orc, nc, ns = db.detect_keypoints_fast(img, ocfg(kw), counts=True)
    assert dc.same_bits(dc.as_rows(got), dc.as_rows(orc))
    assert (prof.candidates, prof.suppressed) == (nc, ns)


def test_detector_large_frame_properties(ctx):
    """1280x720 at the default and a dense setting: equal to the oracle, cap
    per cell held, positions within half a pixel of a FAST corner."""
    from doubleeye_amd import detect, synth
    from oracle import detect_binding as db
    L, _, *_ = synth.slanted_pair(1280, 720, seed=9)
    for kw in ({}, dict(cell=8, per_cell=4, fast_threshold=5)):
        note(f"[detect] 1280x720 {kw}")
        got = detect.detect_keypoints_fast(L, gcfg(kw), ctx=ctx)
        orc = db.detect_keypoints_fast(L, ocfg(kw))
        assert dc.same_bits(dc.as_rows(got), dc.as_rows(orc))
        cell = kw.get("cell", 32)
        cx = (got["x"] + 0.5).astype(int) // cell
        cy = (got["y"] + 0.5).astype(int) // cell
        _, per = np.unique(cy * 10000 + cx, return_counts=True)
        assert per.max() <= kw.get("per_cell", 3)


def test_batch_equals_single(ctx):
    """Images of several sizes and configurations in one device batch give the
    single-call results."""
    from doubleeye_amd import detect, sparse
    sc = dc.scenes()
    imgs = [sc["night_on_L"], dc.image("grid"), sc["morning_on_L"][:200, :300],
            np.full((20, 30), 9, np.uint8), dc.image("noise")]
    kws = [{}, dict(cell=16, per_cell=1), dict(nms_radius=6), {}, dict(fast_threshold=3)]
    frames, want = [], []
    for img, kw in zip(imgs, kws):
        H, W = img.shape
        cap = max(1, detect.max_keypoints(W, H, gcfg(kw)))
        frames.append(detect.DetectFrame(img=ctx.to_device(img),
                                         out=ctx.empty((cap,), sparse.KEYPOINT),
                                         n_out=ctx.empty((1,), np.int32)))
        want.append(detect.detect_keypoints_fast(img, gcfg(kw), ctx=ctx))
    note("[detect] batch of 5")
    detect.detect_keypoints_fast_batch_dev(ctx, frames, [gcfg(k) for k in kws])
    ctx.sync()
    for k, f in enumerate(frames):
        n = int(f.n_out.numpy()[0])
        assert dc.same_bits(dc.as_rows(f.out.numpy()[:n]), dc.as_rows(want[k])), k


def test_census_at_keypoints_equals_oracle(ctx):
    from doubleeye_amd import detect, sparse
    from oracle import detect_binding as db
    rng = np.random.default_rng(33)
    img = dc.image("blobs")
    H, W = img.shape
    kp = np.zeros(500, sparse.KEYPOINT)
    kp["x"] = rng.uniform(-2, W + 2, 500).astype(np.float32)
    kp["y"] = rng.uniform(-2, H + 2, 500).astype(np.float32)
    kp["x"][:20] = np.arange(20) * 0.5          # exact halves round up
    for hw, hh in [(3, 3), (4, 3), (1, 2), (4, 4)]:
        note(f"[detect] census_at {hw}x{hh}")
        got = detect.census_at_keypoints(img, kp, hw, hh, ctx=ctx)
        assert np.array_equal(got, db.census_at_keypoints(img, kp, hw, hh))


def test_refine_disparity_equals_oracle(ctx):
    from doubleeye_amd import detect, sparse
    from oracle import detect_binding as db
    rng = np.random.default_rng(34)
    sc = dc.scenes()
    L, R = sc["evening_on_L"], sc["evening_on_R"]
    n = 3000
    kl = np.zeros(n, sparse.KEYPOINT)
    kl["x"] = rng.uniform(-1, 849, n)
    kl["y"] = rng.uniform(-1, 481, n)
    m = np.zeros(n, sparse.MATCH)
    m["left"] = rng.integers(-1, n + 1, n)
    m["right"] = rng.integers(-1, 3, n)
    m["disparity"] = rng.uniform(-4, 120, n)
    for half in (3, 1, 5):
        note(f"[detect] refine half {half}")
        got = detect.refine_disparity(L, R, kl, kl[:2], m, half, ctx=ctx)
        want = db.refine_disparity(L, R, kl, kl[:2], m, half)
        assert dc.same_bits(got["disparity"], want["disparity"])
        assert np.array_equal(got["left"], m["left"])


@pytest.mark.parametrize("flags", list(dc.PIPE_FLAGS))
def test_pipeline_equals_de_pipe_records(flags, ctx):
    """The device pipeline's DEMR records equal the reference de_pipe's on the
    six scene pairs, and its keypoints and matches equal the oracle's."""
    from doubleeye_amd import detect
    from oracle import detect_binding as db
    det, dr, mc, kw = db.pipe_args(dc.PIPE_FLAGS[flags])
    sc = dc.scenes()
    for s in dc.SCENES:
        note(f"[detect] pipeline {s} {flags}")
        L, R = sc[f"{s}_L"], sc[f"{s}_R"]
        kl, kr, m, rec = detect.sparse_pipeline(L, R, gcfg(det), gcfg(dr), (3, 3),
                                                detect.pipe_match_config(**mc), ctx=ctx, **kw)
        assert dc.same_bits(rec, dc.golden()[f"pipe_{s}_{flags}"]), s
        okl, okr, om, _ = db.pipeline(L, R, ocfg(det), ocfg(dr), (3, 3), db.pipe_match_cfg(**mc),
                                      **kw)
        assert dc.same_bits(dc.as_rows(kl), dc.as_rows(okl))
        assert dc.same_bits(dc.as_rows(kr), dc.as_rows(okr))
        assert np.array_equal(m.view(np.uint32), om.view(np.uint32))


def test_invalid_arguments_raise(ctx):
    from doubleeye_amd import detect
    from doubleeye_amd._lib import DeError
    img = np.zeros((20, 20), np.uint8)
    with pytest.raises(DeError):
        detect.detect_keypoints_fast(img, gcfg(dict(nms_radius=300)), ctx=ctx)
    assert len(detect.detect_keypoints_fast(np.zeros((0, 0), np.uint8), ctx=ctx)) == 0
