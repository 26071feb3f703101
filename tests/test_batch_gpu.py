"""Batched dense MASDA (one launch per stage over uneven pairs), captured
hipGraph plans and the autotuner: every result must equal the per-pair plain
launch bit for bit (same kernels, deterministic reductions), and the plain
launch is held to the oracle by test_dense_gpu.py."""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def note(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def same(a, b):
    a, b = np.asarray(a), np.asarray(b)
    if a.dtype == np.float32:
        return bool(((a.view(np.uint32) == b.view(np.uint32)) | (np.isnan(a) & np.isnan(b))).all())
    return bool((a == b).all())


def make_pairs(ctx, shapes, seed=0):
    from doubleeye_amd import synth
    from doubleeye_amd.dense import Pair
    out = []
    for i, (W, H, D) in enumerate(shapes):
        L, R, _ = synth.pair_for(W, H, D, seed + i, kind="fronto" if i % 2 else "slanted")
        t = dict(L=ctx.to_device(L), R=ctx.to_device(R), disp=ctx.empty((H, W), np.float32),
                 margin=ctx.empty((H, W), np.float32), cand=ctx.empty((6, H, W), np.float32),
                 conf=ctx.empty((H, W), np.uint8), q4=ctx.empty((H, W), np.int16),
                 Lh=L, Rh=R, W=W, H=H, D=D)
        t["pair"] = Pair(t["L"].ptr, t["R"].ptr, W, H, t["disp"].ptr, t["margin"].ptr,
                         t["cand"].ptr, t["conf"].ptr, t["q4"].ptr)
        out.append(t)
    return out


SHAPES = [(320, 240, 64), (424, 240, 96), (640, 360, 128), (200, 120, 192),
          (848, 480, 128), (300, 100, 64), (1280, 720, 192), (424, 240, 64)]


def cfg_for(D, i):
    from doubleeye_amd.dense import Cfg
    return Cfg(dmax=D, iters=(8 if i % 3 else 0), sigma_s=20.0, sigma_r=0.0392,
               min_margin=(0.01 if i == 5 else 0.0), lrc=(i in (2, 6)))


def test_batch_equals_per_pair(ctx):
    items = make_pairs(ctx, SHAPES)
    cfgs = [cfg_for(t["D"], i) for i, t in enumerate(items)]
    note(f"[batch] {len(items)} pairs {SHAPES}")
    ctx.dense_batch_dev([t["pair"] for t in items], cfgs)
    ctx.sync()
    got = [(t["disp"].numpy(), t["margin"].numpy(), t["cand"].numpy(), t["conf"].numpy(),
            t["q4"].numpy()) for t in items]
    for i, t in enumerate(items):
        note(f"[batch] per-pair check {i} {t['W']}x{t['H']}x{t['D']}")
        d, m, c = ctx.run_dense(t["Lh"], t["Rh"], cfgs[i], want_cand=True)
        assert same(got[i][0], d), f"pair {i}: disparity differs from the plain launch"
        assert same(got[i][1], m), f"pair {i}: margin differs"
        assert same(got[i][2], c), f"pair {i}: candidate planes differ"
        # DEDD planes: Q4 = lround(16 d) where d > 0 and finite, else -32768
        q = np.where(np.isfinite(d) & (d > 0), np.floor(np.nan_to_num(d) * np.float32(16) + 0.5),
                     -32768)
        assert same(got[i][4], q.astype(np.int16)), f"pair {i}: Q4 plane differs"
        assert (got[i][3][q == -32768] == 0).all()


def test_plan_graph_replays(ctx):
    shapes = [(424, 240, 64), (320, 240, 96), (640, 360, 128)]
    items = make_pairs(ctx, shapes, seed=40)
    cfgs = [cfg_for(t["D"], i) for i, t in enumerate(items)]
    note("[plan] create")
    plan = ctx.plan([t["pair"] for t in items], cfgs)
    try:
        for rep in range(3):
            # new inputs into the plan's fixed buffers, then replay
            from doubleeye_amd import synth
            for i, t in enumerate(items):
                L, R, _ = synth.pair_for(t["W"], t["H"], t["D"], 100 + rep + i)
                t["L"].upload(L); t["R"].upload(R)
                t["Lh"], t["Rh"] = L, R
            note(f"[plan] replay {rep}")
            plan.launch(ctx.stream)
            ctx.sync()
            for i, t in enumerate(items):
                d, m, _ = ctx.run_dense(t["Lh"], t["Rh"], cfgs[i])
                assert same(t["disp"].numpy(), d), f"replay {rep} pair {i}"
                assert same(t["margin"].numpy(), m)
    finally:
        plan.close()


@pytest.mark.parametrize("w,h,dmax,win,dhi", [(424, 240, 96, (3, 3), 80.0),
                                              (848, 480, 128, (4, 3), 110.0)])
def test_variants_do_not_change_results(ctx, w, h, dmax, win, dhi):
    """Every row-pass and column-pass variant gives the default's planes. At
    848x480 the row waves fit one round: variant 0 keeps forward blocks in LDS
    slots, variant 1 (the frames-in-flight choice) sends them through the
    global scratch, 2 and 3 stream the whole forward row."""
    from doubleeye_amd import synth
    from doubleeye_amd.dense import Cfg
    L, R, _ = synth.slanted_pair(w, h, seed=7, dhi=dhi)
    cfg = Cfg(dmax=dmax, iters=4, sigma_s=20.0, sigma_r=0.0392, census_hw=win[0],
              census_hh=win[1])
    note(f"[variants] {w}x{h} D={dmax} census {2 * win[0] + 1}x{2 * win[1] + 1}")
    ref = ctx.run_dense(L, R, cfg, want_cand=True)
    try:
        for rv in range(4):
            for cv in range(3):
                ctx.set_variants(rv, cv)
                got = ctx.run_dense(L, R, cfg, want_cand=True)
                assert all(same(a, b) for a, b in zip(got, ref)), (rv, cv)
    finally:
        ctx.set_variants(-1, -1)
    note(f"[autotune] {w}x{h} D={dmax}")
    rv, cv = ctx.autotune(w, h, cfg, reps=3)
    assert 0 <= rv < 4 and 0 <= cv < 3
    assert all(same(a, b) for a, b in zip(ctx.run_dense(L, R, cfg, want_cand=True), ref))


C4_SIZES = [(320, 240), (424, 240), (640, 360), (848, 480), (1280, 720)]
C4_DISP = [64, 96, 128, 192]


@pytest.mark.parametrize("parts", [1, 4])
def test_plan_c4_mix_equals_per_pair(ctx, parts, monkeypatch):
    """The 64-pair mix of uneven sizes and disparity ranges (census 9x7, 8
    iterations) as one plan (one graph, or pipelined in 4 parts), launched
    twice with new inputs, against per-pair plain launches."""
    monkeypatch.setenv("DE_PLAN_PARTS", str(parts))
    from doubleeye_amd import synth
    from doubleeye_amd.dense import Cfg, Pair
    rng = np.random.default_rng(4)
    draws = [(C4_SIZES[int(rng.integers(len(C4_SIZES)))], C4_DISP[int(rng.integers(len(C4_DISP)))])
             for _ in range(64)]
    items = []
    for i, ((W, H), D) in enumerate(draws):
        t = dict(L=ctx.empty((H, W), np.uint8), R=ctx.empty((H, W), np.uint8),
                 disp=ctx.empty((H, W), np.float32), margin=ctx.empty((H, W), np.float32),
                 pd=ctx.empty((H, W), np.float32), pm=ctx.empty((H, W), np.float32),
                 W=W, H=H, D=D, cfg=Cfg(dmax=D, iters=8, sigma_s=20.0, sigma_r=0.0392,
                                        census_hw=4, census_hh=3))
        t["pair"] = Pair(t["L"].ptr, t["R"].ptr, W, H, t["disp"].ptr, t["margin"].ptr)
        items.append(t)
    note("[c4] plan of 64 pairs")
    plan = ctx.plan([t["pair"] for t in items], [t["cfg"] for t in items])
    try:
        note(f"[c4] parts {plan.parts}")
        assert plan.parts == parts
        for rep in range(2):
            for i, t in enumerate(items):
                L, R, _ = synth.pair_for(t["W"], t["H"], t["D"], 300 + 64 * rep + i,
                                         kind="fronto" if (i + rep) % 2 else "slanted")
                t["L"].upload(L)
                t["R"].upload(R)
            note(f"[c4] launch {rep}")
            plan.launch(ctx.stream)
            ctx.sync()
            for i, t in enumerate(items):
                ctx.run_dense_dev(t["L"].ptr, t["R"].ptr, t["W"], t["H"], t["cfg"], t["pd"].ptr,
                                  t["pm"].ptr)
            ctx.sync()
            for i, t in enumerate(items):
                assert same(t["disp"].numpy(), t["pd"].numpy()), \
                    f"launch {rep} pair {i} {t['W']}x{t['H']}x{t['D']}: disparity differs"
                assert same(t["margin"].numpy(), t["pm"].numpy()), f"launch {rep} pair {i}: margin"
    finally:
        plan.close()
        for t in items:
            for k in ("L", "R", "disp", "margin", "pd", "pm"):
                t[k].free()
