"""The serving loop (de_server_*): every result served in a mix of streams and
batches equals the plain single-frame call on the same inputs, bit for bit;
the memory budget and slot back-pressure behave as documented."""
import sys
import time

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

DENSE_KW = dict(iters=8, census_hw=4, census_hh=3, sigma_s=20.0, sigma_r=0.0392)


def note(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def dense_plain(ctx, L, R, cfg):
    """disp, q4, conf of a single-pair launch."""
    from doubleeye_amd.dense import Pair
    H, W = L.shape
    dL, dR = ctx.to_device(L), ctx.to_device(R)
    disp = ctx.empty((H, W), np.float32)
    margin = ctx.empty((H, W), np.float32)
    conf = ctx.empty((H, W), np.uint8)
    q4 = ctx.empty((H, W), np.int16)
    ctx.dense_batch_dev([Pair(dL.ptr, dR.ptr, W, H, disp.ptr, margin.ptr, 0, conf.ptr, q4.ptr)],
                        [cfg])
    ctx.sync()
    return disp.numpy(), q4.numpy(), conf.numpy()


def same_dense(a, b):
    return (np.array_equal(a[0].view(np.uint32), b[0].view(np.uint32)) and
            np.array_equal(a[1], b[1]) and np.array_equal(a[2], b[2]))


def make_streams(ctx, srv, spec):
    """spec: list of (kind, W, H, D, n_frames). Returns per-stream dicts with
    frames and their plain results."""
    from doubleeye_amd import knn, synth
    from doubleeye_amd.dense import Cfg
    out = []
    for k, (kind, W, H, D, nf) in enumerate(spec):
        st = dict(kind=kind, W=W, H=H, D=D, frames=[], want=[])
        if kind == "dense":
            st["cfg"] = Cfg(dmax=D, **DENSE_KW)
            st["sid"] = srv.open_dense(W, H, st["cfg"])
        else:
            st["cfg"] = knn.KnnConfig(dmax=D - 1)
            st["sid"] = srv.open_knn(W, H, 2000, st["cfg"])
        for f in range(nf):
            seed = 100 * k + f + 1
            if kind == "dense":
                L, R, _ = synth.slanted_pair(W, H, seed=seed, dhi=min(0.8 * D, 110.0))
                st["frames"].append((L, R, None))
                st["want"].append(dense_plain(ctx, L, R, st["cfg"]))
            else:
                L, R, _ = synth.fronto_pair(W, H, seed)
                rng = np.random.default_rng(seed)
                n = int(rng.integers(200, 2001))
                kp = np.stack([rng.integers(4, W - 4, n), rng.integers(3, H - 3, n)],
                              1).astype(np.int32)
                st["frames"].append((L, R, kp))
                st["want"].append(knn.knn_bp(L, R, kp, st["cfg"], ctx=ctx))
        out.append(st)
    return out


def drain(srv, streams, expect, check=True, timeout_s=60.0):
    """Poll until `expect` results; check each against its plain result."""
    by_sid = {st["sid"]: st for st in streams}
    got, t0 = [], time.time()
    while len(got) < expect:
        assert time.time() - t0 < timeout_s, f"only {len(got)} of {expect} results"
        for r in srv.poll(timeout_ms=50):
            assert r.status == 0, r.status
            st = by_sid[r.stream_id]
            if check:
                res = r.fetch()
                want = st["want"][r.tag % len(st["want"])]
                if st["kind"] == "dense":
                    assert same_dense(res, want), (r.stream_id, r.tag)
                else:
                    assert np.array_equal(res.view(np.uint8), want.view(np.uint8)), \
                        (r.stream_id, r.tag)
            srv.release(r)
            got.append(r)
    return got


def test_mixed_results_equal_plain_calls(ctx):
    from doubleeye_amd.serve import Server
    srv = Server(ctx, hip_streams=3, slots=3)
    try:
        spec = [("dense", 320, 240, 64, 2), ("knn", 424, 240, 96, 2), ("dense", 640, 360, 128, 2),
                ("knn", 848, 480, 128, 2), ("dense", 424, 240, 96, 2), ("knn", 320, 240, 64, 2)]
        streams = make_streams(ctx, srv, spec)
        note("[serve] burst submits")
        pending, got = 0, []
        for rnd in range(6):
            for st in streams:
                L, R, kp = st["frames"][rnd % 2]
                while srv.submit(st["sid"], L, R, kp, tag=rnd) is None:
                    part = drain(srv, streams, 1)
                    got += part
                    pending -= len(part)
                pending += 1
        got += drain(srv, streams, pending)
        assert len(got) == 36
        assert max(r.batch_size for r in got) > 1      # frames were batched together
    finally:
        srv.close()


def test_poisson_arrivals_equal_plain_calls(ctx):
    """8 streams with exponential inter-arrival times (mean 33 ms) for 2 s,
    results collected while frames keep arriving."""
    from doubleeye_amd.serve import Server
    srv = Server(ctx, hip_streams=4, slots=4)
    try:
        spec = [("dense", 320, 240, 64, 2), ("knn", 640, 360, 96, 2), ("dense", 848, 480, 128, 2),
                ("knn", 320, 240, 128, 2), ("dense", 640, 360, 192, 2), ("knn", 848, 480, 64, 2),
                ("dense", 424, 240, 96, 2), ("knn", 1280, 720, 192, 2)]
        streams = make_streams(ctx, srv, spec)
        rng = np.random.default_rng(1)
        t0 = time.perf_counter()
        nxt = [t0 + rng.exponential(0.033) for _ in streams]
        count = [0] * len(streams)
        sent = done = 0
        lat = []
        while time.perf_counter() - t0 < 2.0:
            k = int(np.argmin(nxt))
            while time.perf_counter() < nxt[k]:
                for r in drain(srv, streams, 0):
                    pass
                time.sleep(0.0005)
            st = streams[k]
            L, R, kp = st["frames"][count[k] % 2]
            if srv.submit(st["sid"], L, R, kp, tag=count[k]) is not None:
                sent += 1
            count[k] += 1
            nxt[k] += rng.exponential(0.033)
            res = srv.poll(timeout_ms=0)
            for r in res:
                st2 = [s for s in streams if s["sid"] == r.stream_id][0]
                out = r.fetch()
                want = st2["want"][r.tag % 2]
                if st2["kind"] == "dense":
                    assert same_dense(out, want)
                else:
                    assert np.array_equal(out.view(np.uint8), want.view(np.uint8))
                lat.append(r.latency)
                srv.release(r)
                done += 1
        rest = drain(srv, streams, sent - done)
        lat += [r.latency for r in rest]
        note(f"[serve] {sent} frames, p50 {np.percentile(lat, 50) * 1e3:.2f} ms, "
             f"p99 {np.percentile(lat, 99) * 1e3:.2f} ms")
        assert sent > 100
    finally:
        srv.close()


def test_budget_and_back_pressure(ctx):
    from doubleeye_amd import synth
    from doubleeye_amd._lib import DeError
    from doubleeye_amd.dense import Cfg
    from doubleeye_amd.serve import Server
    cfg = Cfg(dmax=128, **DENSE_KW)
    srv = Server(ctx, mem_budget=600 << 20, hip_streams=1, slots=1)
    try:
        sids = []
        with pytest.raises(DeError):
            while True:
                sids.append(srv.open_dense(848, 480, cfg))
        used, budget = srv.memory()
        assert used <= budget and len(sids) >= 1
        srv.close_stream(sids.pop())
        assert srv.memory()[0] < used
        sid = srv.open_dense(848, 480, cfg)    # the freed budget is reusable
        L, R, _ = synth.slanted_pair(848, 480, seed=3)
        assert srv.submit(sid, L, R, tag=1) is not None
        res = []
        while not res:
            res = srv.poll(timeout_ms=100)
        # the only slot holds an unreleased result: the next frame is refused
        assert srv.submit(sid, L, R, tag=2) is None
        srv.release(res[0])
        assert srv.submit(sid, L, R, tag=3) is not None
        while not srv.poll(timeout_ms=100):
            pass
    finally:
        srv.close()


def test_sixteen_streams_under_load_batch_and_equal_plain_calls(ctx):
    """The C5 shape under load: 16 streams (8 dense, 8 keypoint-graph, sizes
    and disparity ranges from the batching sets) each submitting as fast as
    its slots allow for 3 s, on one HIP lane so that the device is the
    bottleneck. Requests arriving while others run must leave in batches, and
    every result must equal the plain call on its inputs."""
    from doubleeye_amd.serve import Server
    sizes = [(320, 240), (424, 240), (640, 360), (848, 480), (1280, 720)]
    disp = [64, 96, 128, 192]
    rng = np.random.default_rng(16)
    spec = []
    for k in range(16):
        W, H = sizes[int(rng.integers(len(sizes)))]
        D = disp[int(rng.integers(len(disp)))]
        spec.append(("dense" if k % 2 == 0 else "knn", W, H, D, 2))
    srv = Server(ctx, mem_budget=16 << 30, hip_streams=1, slots=4)
    try:
        streams = make_streams(ctx, srv, spec)
        note(f"[serve16] {spec}")
        by_sid = {st["sid"]: st for st in streams}

        def take(timeout_ms):
            out = []
            for r in srv.poll(timeout_ms=timeout_ms):
                assert r.status == 0, r.status
                st = by_sid[r.stream_id]
                res, want = r.fetch(), st["want"][r.tag % 2]
                if st["kind"] == "dense":
                    assert same_dense(res, want), (r.stream_id, r.tag)
                else:
                    assert np.array_equal(res.view(np.uint8), want.view(np.uint8)), \
                        (r.stream_id, r.tag)
                srv.release(r)
                out.append(r)
            return out
        count = [0] * 16
        sent, got = 0, []
        t0 = time.perf_counter()
        while time.perf_counter() - t0 < 3.0:
            for k, st in enumerate(streams):
                L, R, kp = st["frames"][count[k] % 2]
                if srv.submit(st["sid"], L, R, kp, tag=count[k]) is not None:
                    count[k] += 1
                    sent += 1
            got += take(0)
        while len(got) < sent:
            assert time.perf_counter() - t0 < 60.0, f"only {len(got)} of {sent} results"
            got += take(50)
        sizes_seen = [r.batch_size for r in got]
        note(f"[serve16] {sent} frames, mean batch {np.mean(sizes_seen):.2f}, "
             f"max {max(sizes_seen)}")
        assert len(got) == sent and sent > 200
        assert np.mean(sizes_seen) > 1.5 and max(sizes_seen) >= 4
    finally:
        srv.close()
