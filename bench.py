"""Dense MASDA throughput on one MI355X: 848x480 pair, 128 disparities, 8 BP
iterations (the C2 workload), through the C ABI with inputs resident in HBM.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu-baseline]
                    [--frames-in-flight N]

Prints ONE JSON line. `value` = Mpx*disp/s of the whole job (W*H*D per frame
over the timed region). Frames run as a video stream would, several in flight
(--frames-in-flight, default 4: the hardware queues a process gets), each a
captured plan (its own scratch and hipGraph) on its own HIP stream, launched
in turn, so one frame's row pass overlaps another's column pass (the
reference GPU tool pipelines two frames, core/tools/de_dense_cuda.cu:788-841);
every frame runs all four stages and yields the same output. Frames in flight
use the streaming row pass (two waves a SIMD), a lone frame the
register-resident one (see measure_c2).
`single_stream_ms_per_step` is the same run with one frame at a time. `roofline` is for the dominant kernel, timed live with
hipEvents recorded between the stage launches on the launching stream (one
frame at a time, so the kernel runs alone);
`cpu_baseline` times the reference's own CPU tool (oracle/_ref/de_dense,
built from the reference sources) on this host, or the oracle port if that is
absent, on a bounded sample of the same workload.
"""
from __future__ import annotations

import argparse
import json
import os
import re
import subprocess
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

W, H, D = 848, 480, 128
CFG7 = dict(dmax=128, iters=8, sigma_s=20.0, sigma_r=0.0392)
CFG = dict(CFG7, census_hw=4, census_hh=3)                 # C2 as stated: census 9x7
CFG_C1 = dict(dmax=48, iters=5, sigma_s=20.0, sigma_r=0.0392, census_hw=4, census_hh=3)
REF_FLAGS = ["--dmax", "128", "--iters", "8", "--sigma-s", "20", "--sigma-r", "0.0392"]
HBM_PEAK_GBS = 8000.0
ROW_NO_SLOTS = 1           # row-pass launch variant: resident, no LDS forward slots
ROW_STREAM16 = 3           # row-pass launch variant: streaming, backward prefetch 16 blocks
# HBM bytes per launch of each stage kernel at C2 (census 9x7), from rocprofv3
# FETCH_SIZE (x2) and WRITE_SIZE (x1) runs on this workload, the factors read off
# a 256 MiB device copy in the same runs: a quoted profile figure, not measured
# in this run; the source file is named in the output line.
TRAFFIC = {"prep": 15_540_224, "row": 122_910_720, "col": 135_194_624, "solve": 9_889_792}
TRAFFIC_SOURCE = "quoted: profiles/dense_c2_census9x7_848x480x128_kernels.md"


def algorithmic_bytes(w: int, h: int, d: int) -> dict:
    """Algorithmic HBM bytes per frame and per stage kernel.

    Frame: 2WH (pair) + 4WHD (int16 volume written once, read once) + 48WH
    (24 B/px of top-2 candidates written and read) + 8WH (disparity + margin).
    Stages: the row kernel reads the pair and writes the volume once (2WH +
    2WHD); the column kernel reads the volume once and writes 16 B/px candidate
    records (2WHD + 16WH); the solver reads the records and writes disparity and
    margin (16WH + 8WH); prep reads the pair and writes its 28 B/px census and
    coefficient records (2WH + 28WH)."""
    wh = w * h
    return {
        "frame": 2 * wh + 4 * wh * d + 48 * wh + 8 * wh,
        "prep": 2 * wh + 28 * wh,
        "row": 2 * wh + 2 * wh * d,
        "col": 2 * wh * d + 16 * wh,
        "solve": 16 * wh + 8 * wh,
    }


C4_SIZES = [(320, 240), (424, 240), (640, 360), (848, 480), (1280, 720)]
C4_DISP = [64, 96, 128, 192]


def bench_c4(a):
    """C4: 64 uneven pairs (sizes and disparity ranges drawn uniformly from the
    sets above, seed 4), 9x7 census, 8 BP iterations. The batch runs as one
    captured plan (one launch per stage, per-bucket hipGraph); the baseline is
    the same 64 pairs as plain one-pair-per-launch calls. Both must agree bit
    for bit."""
    import numpy as np

    from doubleeye_amd import synth
    from doubleeye_amd.dense import Cfg, Context, Pair

    rng = np.random.default_rng(4)
    draws = [(C4_SIZES[int(rng.integers(len(C4_SIZES)))], C4_DISP[int(rng.integers(len(C4_DISP)))])
             for _ in range(64)]
    ctx = Context(0)
    pool = {}
    items = []
    for i, ((w, h), d) in enumerate(draws):
        if (w, h, d) not in pool:
            pool[(w, h, d)] = synth.pair_for(w, h, d, seed=len(pool) + 11, kind="fronto")[:2]
        L, R = pool[(w, h, d)]
        t = dict(L=ctx.to_device(L), R=ctx.to_device(R), disp=ctx.empty((h, w), np.float32),
                 margin=ctx.empty((h, w), np.float32), W=w, H=h, D=d)
        t["pair"] = Pair(t["L"].ptr, t["R"].ptr, w, h, t["disp"].ptr, t["margin"].ptr)
        t["cfg"] = Cfg(dmax=d, iters=8, sigma_s=20.0, sigma_r=0.0392, census_hw=4, census_hh=3)
        items.append(t)
    pxd = sum(t["W"] * t["H"] * t["D"] for t in items)

    def plain():
        for t in items:
            ctx.run_dense_dev(t["L"].ptr, t["R"].ptr, t["W"], t["H"], t["cfg"], t["disp"].ptr,
                              t["margin"].ptr)

    plain()
    ctx.sync()
    ref = [t["disp"].numpy() for t in items]
    for t in items:
        ctx.fill_dev(t["disp"], 0)
    if a.row_variant >= 0:              # the batch's row-pass variant (plain calls keep theirs)
        ctx.set_variants(a.row_variant, -1)
    plan = ctx.plan([t["pair"] for t in items], [t["cfg"] for t in items])
    ctx.set_variants(-1, -1)
    parts = plan.parts
    plan.launch()
    ctx.sync()
    same = all(np.array_equal(t["disp"].numpy().view(np.uint32), r.view(np.uint32))
               for t, r in zip(items, ref))
    reps = max(2, a.steps // 20)
    for _ in range(2):
        plan.launch()
    ctx.sync()
    t0 = time.perf_counter()
    for _ in range(reps):
        plan.launch()
    ctx.sync()
    ms_batch = (time.perf_counter() - t0) * 1e3 / reps
    plain()
    ctx.sync()
    t0 = time.perf_counter()
    for _ in range(reps):
        plain()
    ctx.sync()
    ms_plain = (time.perf_counter() - t0) * 1e3 / reps
    plan.close()
    out = {
        "metric": "batched dense Mpx*disp/s, 64 uneven pairs (C4)",
        "value": round(pxd / (ms_batch * 1e-3) / 1e6, 1), "unit": "Mpx*disp/s",
        "n_gpus": 1, "steps": reps, "warmup": 2, "ms_per_step": round(ms_batch, 3),
        "frames_per_s": round(64e3 / ms_batch, 1),
        "plain_ms_per_batch": round(ms_plain, 3),
        "plain_value": round(pxd / (ms_plain * 1e-3) / 1e6, 1),
        "batched_over_plain": round(ms_plain / ms_batch, 3),
        "plan_parts": parts,
        "identical_to_plain": same,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "i16 volume / f32 messages",
        "data": "synthetic (fronto-plane pairs, doubleeye_amd.synth, draw seed 4)",
        "config": {"workload": "64 pairs, sizes {320x240,424x240,640x360,848x480,1280x720}, "
                               "D {64,96,128,192}, census 9x7, 8 BP iterations",
                   "px_disp_per_batch": pxd},
    }
    ctx.close()
    print(json.dumps(out), flush=True)


def cpu_baseline(L, R, budget_s: float = 12.0):
    """Reference CPU tool (census 7x7, its only window) at --threads 1 on the
    same pair, repeated to ~budget_s, and the same tool on up to 16 threads for
    half that time (reported beside it). Falls back to the oracle port at 7x7
    where the reference build is absent."""
    ref = os.path.join(ROOT, "oracle", "_ref", "de_dense")
    if not os.access(ref, os.X_OK):
        return cpu_baseline_port(L, R, budget_s, CFG7)
    with tempfile.TemporaryDirectory() as td:
        lp, rp = os.path.join(td, "l.y8"), os.path.join(td, "r.y8")
        L.tofile(lp)
        R.tofile(rp)
        times = []
        t_end = time.time() + budget_s
        while time.time() < t_end or not times:
            out = subprocess.run([ref, lp, rp, str(W), str(H), "--threads", "1", *REF_FLAGS],
                                 capture_output=True, text=True, check=True).stdout
            m = re.findall(r"total ([0-9.]+) ms", out)
            times.append(float(m[-1]))
        ms = sum(times) / len(times)
        # the same tool on the host's share of cores (16 per GPU here), for context
        nth = max(1, min(16, os.cpu_count() or 1))
        mt = []
        t_end = time.time() + budget_s / 2
        while time.time() < t_end or not mt:
            out = subprocess.run([ref, lp, rp, str(W), str(H), "--threads", str(nth),
                                  *REF_FLAGS], capture_output=True, text=True,
                                 check=True).stdout
            mt.append(float(re.findall(r"total ([0-9.]+) ms", out)[-1]))
        ms_mt = sum(mt) / len(mt)
    return {"value": round(W * H * D / (ms * 1e-3) / 1e6, 2), "unit": "Mpx*disp/s",
            "cores": 1, "kind": "reference",
            "sample": f"{len(times)} frames of 848x480x128 iters 8 census 7x7, de_dense "
                      f"--threads 1 (its own census+cost+solve timer), mean {ms:.1f} ms/frame",
            "all_threads": {"value": round(W * H * D / (ms_mt * 1e-3) / 1e6, 2),
                            "threads": nth, "ms_per_frame": round(ms_mt, 1),
                            "frames": len(mt)}}


def cpu_baseline_port(L, R, budget_s: float = 12.0, cfg: dict | None = None, w: int = W,
                      h: int = H):
    """The oracle's restatement of the reference pipeline (single thread, the
    reference's arithmetic) on the same pair, repeated to ~budget_s. The
    reference's own tool has no 9x7 census, so the stated C2 workload is timed
    through the port."""
    from oracle import binding as orc
    cfg = CFG if cfg is None else cfg
    d = cfg["dmax"] - cfg.get("dmin", 1) + 1
    c = orc.cfg(**cfg)
    times = []
    t_end = time.time() + budget_s
    while time.time() < t_end or not times:
        t0 = time.perf_counter()
        orc.dense_run(L, R, c)
        times.append((time.perf_counter() - t0) * 1e3)
    ms = sum(times) / len(times)
    win = f"{2 * cfg.get('census_hw', 3) + 1}x{2 * cfg.get('census_hh', 3) + 1}"
    return {"value": round(w * h * d / (ms * 1e-3) / 1e6, 2), "unit": "Mpx*disp/s",
            "cores": 1, "kind": "port",
            "sample": f"{len(times)} frames of {w}x{h}x{d} iters {cfg['iters']} census {win} "
                      f"through the oracle (C++ restatement of de_dense --threads 1), "
                      f"mean {ms:.1f} ms/frame"}


def bench_sparse(a):
    """Sparse keypoint MASDA (match_masda): 2000 left keypoints on the 848x480
    random-dot pair, 8 candidates, 10 iterations, inputs resident in HBM, one
    frame per launch sequence and a batch of 16 frames. The CPU baseline is the
    reference's own match_masda (oracle/_ref/libref_sparse.so, built from
    core/src/match.cpp), single thread, on the same frame."""
    import numpy as np

    from doubleeye_amd import sparse, synth
    from doubleeye_amd.dense import Context

    ctx = Context(0)
    cfg = sparse.MatchConfig(max_candidates=8, iterations=10)
    probs = [synth.sparse_problem(2000, 848, 480, seed) for seed in range(3, 3 + 16)]

    def frame(p):
        lxy, dl, rxy, dr, _ = p
        L, R = sparse.keypoints(lxy), sparse.keypoints(rxy)
        return sparse.SparseFrame(ctx.to_device(L), ctx.to_device(dl), ctx.to_device(R),
                                  ctx.to_device(dr), ctx.empty((min(len(L), len(R)),), sparse.MATCH),
                                  ctx.empty((1,), np.int32), ctx.empty((1,), sparse.STATS))

    frames = [frame(p) for p in probs]
    one = frames[:1]
    want = sparse.match_masda(*probs[0][:4], cfg, ctx=ctx)
    sparse.match_masda_batch_dev(ctx, one, [cfg])
    ctx.sync()
    n0 = int(one[0].n_out.numpy()[0])
    same = bool(n0 == len(want) and np.array_equal(one[0].out.numpy()[:n0].view(np.uint8),
                                                   want.view(np.uint8)))
    edges = int(one[0].stats.numpy()[0]["candidates"])
    reps = max(20, a.steps)
    for _ in range(5):
        sparse.match_masda_batch_dev(ctx, one, [cfg])
    ctx.sync()
    t0 = time.perf_counter()
    for _ in range(reps):
        sparse.match_masda_batch_dev(ctx, one, [cfg])
    ctx.sync()
    ms_one = (time.perf_counter() - t0) * 1e3 / reps
    sparse.match_masda_batch_dev(ctx, frames, [cfg] * len(frames))
    ctx.sync()
    t0 = time.perf_counter()
    for _ in range(reps):
        sparse.match_masda_batch_dev(ctx, frames, [cfg] * len(frames))
    ctx.sync()
    ms_batch = (time.perf_counter() - t0) * 1e3 / reps
    out = {
        "metric": "sparse match_masda frames/s, 2000 keypoints (848x480)",
        "value": round(1e3 / ms_one, 1), "unit": "frames/s",
        "n_gpus": 1, "steps": reps, "warmup": 5, "ms_per_step": round(ms_one, 4),
        "edges": edges, "edges_per_s": round(edges / (ms_one * 1e-3), 1),
        "batch16_ms": round(ms_batch, 4), "batch16_frames_per_s": round(16e3 / ms_batch, 1),
        "identical_to_single_call": same,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32 messages / u64 census", "data": "synthetic (doubleeye_amd.synth.sparse_problem)",
        "config": {"workload": "match_masda, 2000 left kp, max_candidates 8, 10 iterations"},
    }
    if not a.no_cpu_baseline:
        try:
            sys.path.insert(0, ROOT)
            from oracle import sparse_binding as sb
            c = sb.match_cfg(max_candidates=8, iterations=10)
            lxy, dl, rxy, dr, _ = probs[0]
            sb.ref_match_masda(lxy, dl, rxy, dr, c)
            n, t0 = 0, time.perf_counter()
            while time.perf_counter() - t0 < min(a.cpu_budget, 5.0):
                sb.ref_match_masda(lxy, dl, rxy, dr, c)
                n += 1
            ms = (time.perf_counter() - t0) * 1e3 / n
            out["cpu_baseline"] = {"value": round(1e3 / ms, 1), "unit": "frames/s", "cores": 1,
                                   "kind": "reference", "ms_per_frame": round(ms, 3),
                                   "sample": f"{n} runs of the same frame"}
        except Exception as e:   # reference harness absent on this host
            out["cpu_baseline"] = {"error": str(e)}
    ctx.close()
    print(json.dumps(out), flush=True)


def bench_c3(a):
    """C3: keypoint-graph BP. 2000 keypoints drawn uniformly (seeded) on the
    848x480 random-dot pair, 9x7 census, 8 labels from d in 0..127 by Hamming
    distance, k = 8 nearest neighbours, 10 min-sum iterations with the
    truncated-linear term (trunc 6 px), top-2 per keypoint. Inputs resident in
    HBM; one frame per launch sequence, and a batch of 16 frames. The CPU
    baseline is the oracle's restatement (the reference has no such solver),
    single thread."""
    import numpy as np

    from doubleeye_amd import knn, synth
    from doubleeye_amd.dense import Context

    ctx = Context(0)
    cfg = knn.KnnConfig()
    W_, H_, N_ = 848, 480, 2000

    def problem(seed):
        L, R, _ = synth.fronto_pair(W_, H_, seed)
        rng = np.random.default_rng(seed + 16)
        kp = np.stack([rng.integers(4, W_ - 4, N_), rng.integers(3, H_ - 3, N_)], 1).astype(np.int32)
        return L, R, kp

    probs = [problem(s) for s in range(1, 17)]
    frames = [knn.KnnFrame(ctx.to_device(L), ctx.to_device(R), ctx.to_device(kp),
                           ctx.empty((N_,), knn.OUT)) for L, R, kp in probs]
    one = frames[:1]
    want = knn.knn_bp(*probs[0], cfg, ctx=ctx)
    knn.knn_bp_batch_dev(ctx, one, [cfg])
    ctx.sync()
    got = one[0].out.numpy()
    same = bool(np.array_equal(got.view(np.uint8), want.view(np.uint8)))
    reps = max(20, a.steps)
    for _ in range(5):
        knn.knn_bp_batch_dev(ctx, one, [cfg])
    ctx.sync()
    t0 = time.perf_counter()
    for _ in range(reps):
        knn.knn_bp_batch_dev(ctx, one, [cfg])
    ctx.sync()
    ms_one = (time.perf_counter() - t0) * 1e3 / reps
    knn.knn_bp_batch_dev(ctx, frames, [cfg] * 16)
    ctx.sync()
    t0 = time.perf_counter()
    for _ in range(reps):
        knn.knn_bp_batch_dev(ctx, frames, [cfg] * 16)
    ctx.sync()
    ms_batch = (time.perf_counter() - t0) * 1e3 / reps
    gt = synth.fronto_pair(W_, H_, 1)[2]
    kp0 = probs[0][2]
    acc = float((got["d1"] == gt[kp0[:, 1], kp0[:, 0]]).mean())
    out = {
        "metric": "keypoint-graph BP frames/s, 2000 keypoints x 128 disparities (C3)",
        "value": round(1e3 / ms_one, 1), "unit": "frames/s",
        "n_gpus": 1, "steps": reps, "warmup": 5, "ms_per_step": round(ms_one, 4),
        "batch16_ms": round(ms_batch, 4), "batch16_frames_per_s": round(16e3 / ms_batch, 1),
        "identical_to_single_call": same, "d1_accuracy_vs_truth": round(acc, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "u64 census / f32 messages",
        "data": "synthetic (fronto random-dot pair, uniform keypoints, seeds 1..16)",
        "config": {"workload": "2000 kp, 848x480, census 9x7, 8 labels of d 0..127, k=8 NN, "
                               "10 iterations, trunc 6"},
    }
    if not a.no_cpu_baseline:
        sys.path.insert(0, ROOT)
        from oracle import knn_binding as kb
        L, R, kp = probs[0]
        c = kb.cfg()
        n, t0 = 0, time.perf_counter()
        while time.perf_counter() - t0 < min(a.cpu_budget, 10.0):
            kb.run(L, R, kp, c)
            n += 1
        ms = (time.perf_counter() - t0) * 1e3 / n
        out["cpu_baseline"] = {"value": round(1e3 / ms, 2), "unit": "frames/s", "cores": 1,
                               "kind": "port", "ms_per_frame": round(ms, 2),
                               "sample": f"{n} runs of frame 1 (oracle restatement; the "
                                         "reference has no k-NN BP)"}
    ctx.close()
    print(json.dumps(out), flush=True)


def bench_c5(a):
    print(json.dumps(run_c5(a.duration, a.rate)), flush=True)


def run_c5(duration: float, rate_hz: float = 30.0) -> dict:
    """C5: serving loop. 16 camera streams, each rate_hz (30 Hz) with Poisson
    arrivals (exponential inter-arrival, mean 33.3 ms), alternately dense and
    keypoint-graph, sizes and disparity ranges drawn from the C4 sets with a
    per-stream seed; dense frames from the slanted-plane generator (census
    9x7, 8 iterations), keypoint frames from the fronto generator (2000
    uniform keypoints). Per-stream buffers and scratch resident in HBM,
    admission by the W x H x D memory budget, continuous batching over 4 HIP
    streams. Latency = submit to results in host memory."""
    import numpy as np

    from doubleeye_amd import knn, synth
    from doubleeye_amd.dense import Cfg, Context
    from doubleeye_amd.serve import Server

    ctx = Context(0)
    srv = Server(ctx, mem_budget=16 << 30, hip_streams=4, slots=4)
    streams = []
    for k in range(16):
        rng = np.random.default_rng(500 + k)
        W_, H_ = C4_SIZES[int(rng.integers(len(C4_SIZES)))]
        D_ = C4_DISP[int(rng.integers(len(C4_DISP)))]
        st = dict(kind="dense" if k % 2 == 0 else "knn", W=W_, H=H_, D=D_, frames=[])
        if st["kind"] == "dense":
            st["sid"] = srv.open_dense(W_, H_, Cfg(dmax=D_, iters=8, census_hw=4, census_hh=3,
                                                   sigma_s=20.0, sigma_r=0.0392))
            for f in range(2):
                L, R, _ = synth.slanted_pair(W_, H_, seed=1000 * k + f, dhi=min(0.8 * D_, 110.0))
                st["frames"].append((L, R, None))
        else:
            st["sid"] = srv.open_knn(W_, H_, 2000, knn.KnnConfig(dmax=D_ - 1))
            for f in range(2):
                L, R, _ = synth.fronto_pair(W_, H_, 1000 * k + f)
                kp = np.stack([rng.integers(4, W_ - 4, 2000), rng.integers(3, H_ - 3, 2000)],
                              1).astype(np.int32)
                st["frames"].append((L, R, kp))
        streams.append(st)
    by_sid = {st["sid"]: st for st in streams}
    # first result of every (stream, frame) is kept to check later ones against
    first = {}
    lat = {"dense": [], "knn": []}
    batches = []
    mismatches = 0
    done = 0

    def collect(timeout_ms):
        nonlocal mismatches, done
        for r in srv.poll(timeout_ms=timeout_ms):
            st = by_sid[r.stream_id]
            if r.status != 0:
                raise RuntimeError(f"frame failed: status {r.status}")
            key = (r.stream_id, r.tag % 2)
            if key not in first or (r.tag // 2) % 25 == 0:   # check a sample
                out = r.fetch()
                if key not in first:
                    first[key] = out
                else:
                    ref = first[key]
                    if st["kind"] == "dense":
                        ok = all(np.array_equal(x.view(np.uint8), y.view(np.uint8))
                                 for x, y in zip(out, ref))
                    else:
                        ok = np.array_equal(out.view(np.uint8), ref.view(np.uint8))
                    mismatches += 0 if ok else 1
            lat[st["kind"]].append(r.latency)
            batches.append(r.batch_size)
            srv.release(r)
            done += 1

    rng = np.random.default_rng(5)
    count = [0] * 16
    dropped = 0
    warm = 1.0
    t0 = time.perf_counter()
    nxt = [t0 + rng.exponential(1 / rate_hz) for _ in range(16)]
    submitted_timed = 0
    pxd = 0
    last_note = t0
    while True:
        if time.perf_counter() - last_note > 10.0:
            last_note = time.perf_counter()
            print(f"[c5] {last_note - t0:.0f} s, {done} frames done", file=sys.stderr, flush=True)
        k = int(np.argmin(nxt))
        if nxt[k] - t0 > duration + warm:
            break
        while True:
            dt = nxt[k] - time.perf_counter()
            if dt <= 0:
                break
            collect(max(0, min(int(dt * 1e3), 2)))
        st = streams[k]
        L, R, kp = st["frames"][count[k] % 2]
        timed = nxt[k] - t0 >= warm
        if srv.submit(st["sid"], L, R, kp, tag=count[k]) is None:
            dropped += 1
        elif timed:
            submitted_timed += 1
            if st["kind"] == "dense":
                pxd += st["W"] * st["H"] * st["D"]
        if not timed:
            lat["dense"].clear(); lat["knn"].clear(); batches.clear()
        count[k] += 1
        nxt[k] += rng.exponential(1 / rate_hz)
    t_end = time.perf_counter()
    while done < sum(count) - dropped:
        collect(50)
    wall = t_end - t0 - warm
    alll = lat["dense"] + lat["knn"]
    pct = lambda xs, p: round(float(np.percentile(xs, p)) * 1e3, 3) if xs else None
    out = {
        "metric": f"serving p50/p99 frame latency, 16 mixed camera streams at {rate_hz:g} Hz (C5)",
        "value": pct(alll, 99), "unit": "ms (p99)", "p50_ms": pct(alll, 50),
        "p99_ms": pct(alll, 99), "max_ms": pct(alll, 100),
        "dense_p50_ms": pct(lat["dense"], 50), "dense_p99_ms": pct(lat["dense"], 99),
        "knn_p50_ms": pct(lat["knn"], 50), "knn_p99_ms": pct(lat["knn"], 99),
        "frames_per_s": round(submitted_timed / wall, 1),
        "dense_mpx_disp_per_s": round(pxd / wall / 1e6, 1),
        "frames": submitted_timed, "dropped_busy": dropped,
        "mean_batch": round(float(np.mean(batches)), 2) if batches else None,
        "max_batch": int(max(batches)) if batches else None,
        "sampled_results_identical": mismatches == 0,
        "n_gpus": 1, "steps": submitted_timed, "warmup": int(warm),
        "ms_per_step": round(wall * 1e3 / max(1, submitted_timed), 4),
        "higher_is_better": False, "scaling": "weak", "vs_baseline": None,
        "dtype": "i16 volume / f32 messages",
        "data": "synthetic (slanted and fronto random-dot pairs, per-stream seeds)",
        "offered_frames_per_s": round(16 * rate_hz, 1),
        "config": {"workload": f"16 streams (8 dense, 8 keypoint-graph), Poisson {rate_hz:g} Hz, "
                               f"{duration:.0f} s, sizes/D from the C4 sets",
                   "streams": [f"{st['kind']} {st['W']}x{st['H']} D{st['D']}" for st in streams]},
    }
    srv.close()
    ctx.close()
    return out


def bench_dense_masda(a):
    """The NumPy dense MASDA of the reference's article (dense_masda.run,
    dense_masda.py:82-110) at its Middlebury size: 450x375, d 1..60, census 7x7,
    12 iterations, lambda = gamma = -0.1 (10.1 M edges). The device pipeline
    (edge list, float64 messages, greedy decision, disparity map) from host
    images to a host disparity map. The CPU baseline is the oracle's NumPy
    restatement of the same functions, single thread, one run."""
    import numpy as np

    from doubleeye_amd import masda_stereo as gm
    from doubleeye_amd import synth
    from doubleeye_amd.dense import Context

    ctx = Context(0)
    W_, H_ = 450, 375
    L, R, gt = synth.fronto_pair(W_, H_, 9)
    disp, hist = gm.dense_masda(L, R, 1, 60, iters=12, ctx=ctx)
    reps = max(3, a.steps // 40)
    t0 = time.perf_counter()
    for _ in range(reps):
        disp, hist = gm.dense_masda(L, R, 1, 60, iters=12, ctx=ctx)
    ms_ = (time.perf_counter() - t0) * 1e3 / reps
    E = 375 - 6
    edges = sum(max(0, 450 - 6 - d) for d in range(1, 61)) * E
    ok = np.isfinite(disp)
    out = {
        "metric": "NumPy-shape dense MASDA (dense_masda.run) time, 450x375, d 1..60, 12 it",
        "value": round(ms_, 3), "unit": "ms", "edges": edges,
        "edges_per_s": round(edges * len(hist) / (ms_ * 1e-3), 1),
        "iterations_run": len(hist), "coverage": round(float(ok.mean()), 4),
        "bad_1px_vs_truth": round(float((np.abs(disp[ok] - gt[ok]) > 1.0).mean()), 4),
        "n_gpus": 1, "steps": reps, "warmup": 1, "ms_per_step": round(ms_, 3),
        "higher_is_better": False, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64 messages", "data": "synthetic (fronto random-dot pair 450x375, seed 9)",
        "config": {"workload": "dense_problem + masda_sparse + disparity map, host to host"},
    }
    if not a.no_cpu_baseline:
        sys.path.insert(0, ROOT)
        from oracle import masda_np as mn
        t0 = time.perf_counter()
        ei, ej, se, n = mn.dense_problem(L, R, 1, 60)
        oa, oh = mn.masda_sparse(ei, ej, se, n, n, -0.1, -0.1, iters=12)
        t = time.perf_counter() - t0
        out["cpu_baseline"] = {"value": round(t * 1e3, 1), "unit": "ms", "cores": 1,
                               "kind": "port", "sample": "one run of the same frame (NumPy "
                               "restatement of dense_problem + masda_sparse)",
                               "same_history": bool(np.array_equal(oh, hist))}
    ctx.close()
    print(json.dumps(out), flush=True)


def _scene_pairs():
    import numpy as np
    z = np.load(os.path.join(ROOT, "tests", "golden", "scenes.npz"))
    names = ["night_on", "night_off", "evening_on", "evening_off", "morning_on", "morning_off"]
    return names, [(z[f"{s}_L"], z[f"{s}_R"]) for s in names]


def bench_detect(a):
    """detect_keypoints_fast (DetectorConfig defaults) on the reference's six
    848x480 IR scene pairs (12 images), images resident in HBM: one image per
    launch chain, and all 12 in one batch. Per-stage times come from the host
    entry's DetectProfile on the first image. The CPU baseline is the
    reference's own detect_keypoints_fast (oracle/_ref/libref_sparse.so, built
    from core/src/preproc.cpp), single thread, on the same 12 images."""
    import numpy as np

    from doubleeye_amd import detect, sparse
    from doubleeye_amd.dense import Context

    ctx = Context(0)
    _, pairs = _scene_pairs()
    imgs = [im for p in pairs for im in p]
    cfg = detect.DetectorConfig()
    cap = detect.max_keypoints(848, 480, cfg)
    frames = [detect.DetectFrame(ctx.to_device(im), ctx.empty((cap,), sparse.KEYPOINT),
                                 ctx.empty((1,), np.int32)) for im in imgs]
    prof = detect.DetectProfile()
    want = [detect.detect_keypoints_fast(im, cfg, prof if k == 0 else None, ctx=ctx)
            for k, im in enumerate(imgs)]
    reps = max(20, a.steps)
    for f in frames[:1] * 5:
        detect.detect_keypoints_fast_batch_dev(ctx, [f], [cfg])
    ctx.sync()
    t0 = time.perf_counter()
    for k in range(reps):
        detect.detect_keypoints_fast_batch_dev(ctx, [frames[k % 12]], [cfg])
    ctx.sync()
    ms_one = (time.perf_counter() - t0) * 1e3 / reps
    detect.detect_keypoints_fast_batch_dev(ctx, frames, [cfg] * 12)
    ctx.sync()
    same = all(np.array_equal(f.out.numpy()[:int(f.n_out.numpy()[0])].view(np.uint32),
                              w.view(np.uint32)) for f, w in zip(frames, want))
    t0 = time.perf_counter()
    for _ in range(reps):
        detect.detect_keypoints_fast_batch_dev(ctx, frames, [cfg] * 12)
    ctx.sync()
    ms_batch = (time.perf_counter() - t0) * 1e3 / reps
    out = {
        "metric": "detect_keypoints_fast images/s (848x480 IR scenes, default config)",
        "value": round(1e3 / ms_one, 1), "unit": "images/s",
        "n_gpus": 1, "steps": reps, "warmup": 5, "ms_per_step": round(ms_one, 4),
        "batch12_ms": round(ms_batch, 4), "batch12_images_per_s": round(12e3 / ms_batch, 1),
        "keypoints_per_image": round(float(np.mean([len(w) for w in want])), 1),
        "stage_ms_first_image": {"fast": round(prof.fast_ms, 4), "nms": round(prof.nms_ms, 4),
                                 "refine": round(prof.refine_ms, 4)},
        "candidates_first_image": prof.candidates, "suppressed_first_image": prof.suppressed,
        "batch_identical_to_single_calls": same,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "u8 pixels / f32 response / f64 texture", "data": "the reference's IR scenes",
        "config": {"workload": "detect_keypoints_fast, 848x480, DetectorConfig defaults"},
    }
    if not a.no_cpu_baseline:
        try:
            sys.path.insert(0, ROOT)
            from oracle import detect_binding as db
            c = db.detector_cfg()
            n, t0 = 0, time.perf_counter()
            while time.perf_counter() - t0 < min(a.cpu_budget, 5.0):
                db.ref_detect_keypoints_fast(imgs[n % 12], c)
                n += 1
            ms = (time.perf_counter() - t0) * 1e3 / n
            out["cpu_baseline"] = {"value": round(1e3 / ms, 1), "unit": "images/s", "cores": 1,
                                   "kind": "reference", "ms_per_image": round(ms, 3),
                                   "sample": f"{n} images cycling over the 12"}
        except Exception as e:   # reference harness absent on this host
            out["cpu_baseline"] = {"error": str(e)}
    ctx.close()
    print(json.dumps(out), flush=True)


def bench_pipe(a):
    """core/tools/de_pipe.cpp's per-pair work (detect both images, census_at,
    match_masda with de_pipe's settings, refine_disparity) on the six IR scene
    pairs, host images in, DEMR records out (de_sparse_pipeline). The CPU
    baseline is the reference's own de_pipe binary (built from its sources by
    oracle/Makefile.ref) fed the same pairs as DEIR packets, single thread."""
    import struct

    import numpy as np

    from doubleeye_amd import detect
    from doubleeye_amd.dense import Context

    ctx = Context(0)
    names, pairs = _scene_pairs()
    mcfg = detect.pipe_match_config()
    for L, R in pairs:
        detect.sparse_pipeline(L, R, cfg=mcfg, ctx=ctx)
    reps = max(12, a.steps // 2)
    nrec = 0
    t0 = time.perf_counter()
    for k in range(reps):
        L, R = pairs[k % 6]
        nrec += len(detect.sparse_pipeline(L, R, cfg=mcfg, ctx=ctx)[3])
    ms_ = (time.perf_counter() - t0) * 1e3 / reps
    out = {
        "metric": "de_pipe sparse stereo pairs/s (848x480 IR scenes, host to host)",
        "value": round(1e3 / ms_, 1), "unit": "pairs/s",
        "n_gpus": 1, "steps": reps, "warmup": 6, "ms_per_step": round(ms_, 4),
        "matches_per_pair": round(nrec / reps, 1),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "u8 pixels / f32 messages", "data": "the reference's IR scenes",
        "config": {"workload": "detect x2 + census_at + match_masda (de_pipe settings) + "
                               "refine_disparity, 848x480"},
    }
    if not a.no_cpu_baseline:
        exe = os.path.join(ROOT, "oracle", "_ref", "de_pipe")
        if os.path.exists(exe):
            k = 0
            chunks = []
            for rep in range(8):
                for L, R in pairs:
                    for img, st in ((L, 1), (R, 0)):
                        chunks.append(b"DEIR" + struct.pack("<HHBBI", 848, 480, st, 0, k)
                                      + np.ascontiguousarray(img).tobytes())
                    k += 1
            data = b"".join(chunks)
            t0 = time.perf_counter()
            subprocess.run([exe], input=data, capture_output=True, check=True)
            t = time.perf_counter() - t0
            out["cpu_baseline"] = {"value": round(k / t, 1), "unit": "pairs/s", "cores": 1,
                                   "kind": "reference", "ms_per_pair": round(t * 1e3 / k, 3),
                                   "sample": f"{k} pairs (the six scenes x 8) through de_pipe"}
        else:
            out["cpu_baseline"] = {"error": "oracle/_ref/de_pipe not built on this host"}
    ctx.close()
    print(json.dumps(out), flush=True)


def bench_full(a):
    """The FP32 full-volume path (`de_dense --topk 0`: the float volume and the
    full-volume row solver) on the C2 pair, 848x480, d 1..128, 8 iterations,
    inputs resident in HBM, one frame at a time. The host-pointer call must give
    the same disparities. CPU baseline: the reference `de_dense --topk 0
    --threads 1` on the same pair, one run (its own census+cost+solve timer)."""
    import numpy as np

    from doubleeye_amd import synth
    from doubleeye_amd.dense import Cfg, Context

    L, R, _ = synth.slanted_pair(W, H, seed=2)
    ctx = Context(0)
    cfg = Cfg(**CFG)
    dL, dR = ctx.to_device(L), ctx.to_device(R)
    dd, dm = ctx.empty((H, W), np.float32), ctx.empty((H, W), np.float32)

    def step():
        ctx.run_dense_full_dev(dL.ptr, dR.ptr, W, H, cfg, dd.ptr, dm.ptr)

    step()
    ctx.sync()
    first = dd.numpy()
    reps = max(10, a.steps // 4)
    for _ in range(max(2, a.warmup // 4)):
        step()
    ctx.sync()
    t0 = time.perf_counter()
    for _ in range(reps):
        step()
    ctx.sync()
    ms = (time.perf_counter() - t0) * 1e3 / reps
    host_d, _ = ctx.run_dense_full(L, R, cfg)
    same = bool(np.array_equal(host_d.view(np.uint32), first.view(np.uint32)) and
                np.array_equal(dd.numpy().view(np.uint32), first.view(np.uint32)))
    out = {
        "metric": "dense Mpx*disp/s, FP32 full-volume path (--topk 0) at 848x480x128",
        "value": round(W * H * D / (ms * 1e-3) / 1e6, 1), "unit": "Mpx*disp/s",
        "n_gpus": 1, "steps": reps, "warmup": max(2, a.warmup // 4), "ms_per_step": round(ms, 4),
        "deterministic_and_equal_to_host_call": same,
        "coverage": round(float(np.isfinite(first).mean()), 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32 volume / f32 messages",
        "data": "synthetic (seeded slanted-plane pair, doubleeye_amd.synth.slanted_pair seed 2)",
        "config": {"workload": "de_dense --topk 0, 848x480, d 1..128, 8 iterations"},
    }
    exe = os.path.join(ROOT, "oracle", "_ref", "de_dense")
    if not a.no_cpu_baseline and os.access(exe, os.X_OK):
        with tempfile.TemporaryDirectory() as td:
            lp, rp = os.path.join(td, "l.y8"), os.path.join(td, "r.y8")
            L.tofile(lp)
            R.tofile(rp)
            res = subprocess.run([exe, lp, rp, str(W), str(H), "--threads", "1", "--topk", "0",
                                  *REF_FLAGS], capture_output=True, text=True, check=True).stdout
            t = float(re.findall(r"total ([0-9.]+) ms", res)[-1])
        out["cpu_baseline"] = {"value": round(W * H * D / (t * 1e-3) / 1e6, 2),
                               "unit": "Mpx*disp/s", "cores": 1, "kind": "reference",
                               "sample": f"one frame of de_dense --topk 0 --threads 1, {t:.0f} ms "
                                         "(its own census+cost+solve timer)"}
    ctx.close()
    print(json.dumps(out), flush=True)


def bench_host(a):
    """Host-to-host C2 (848x480, d 1..128, census 9x7, 8 iterations): the pair
    in host memory, disparity and margin back in host memory, one frame a
    call, against the resident call (pair and outputs in HBM, a stream wait
    after each call). Host legs: pageable numpy arrays (fresh outputs each
    call, and reused ones), de_host_alloc (pinned) arrays, and the DEIR -> DEDD
    stream front end (Q4 disparity and confidence out) fed one pair a call and
    a whole run of pairs in one call (pairs in flight)."""
    import numpy as np

    from doubleeye_amd import synth
    from doubleeye_amd.dense import Cfg, Context, host_empty
    from doubleeye_amd.stream import Stream, deir_packet, parse_dedd

    ctx = Context(0)
    L, R, _ = synth.slanted_pair(W, H, seed=2)
    cfg = Cfg(**CFG)
    reps = max(10, a.steps // 4)

    def timed(fn):
        for _ in range(3):
            out = fn()
        t0 = time.perf_counter()
        for _ in range(reps):
            out = fn()
        return (time.perf_counter() - t0) * 1e3 / reps, out

    dL, dR = ctx.to_device(L), ctx.to_device(R)
    dd, dm = ctx.empty((H, W), np.float32), ctx.empty((H, W), np.float32)

    def res():
        ctx.run_dense_dev(dL.ptr, dR.ptr, W, H, cfg, dd.ptr, dm.ptr)
        ctx.sync()
    ms_res = timed(res)[0]
    want = dd.numpy()
    ms_fresh, o1 = timed(lambda: ctx.run_dense(L, R, cfg))
    reuse = (np.empty((H, W), np.float32), np.empty((H, W), np.float32))
    ms_reuse, o2 = timed(lambda: ctx.run_dense(L, R, cfg, out=reuse))
    pL, pR = host_empty((H, W), np.uint8), host_empty((H, W), np.uint8)
    pL[...] = L
    pR[...] = R
    pin = (host_empty((H, W), np.float32), host_empty((H, W), np.float32))
    ms_pin, o3 = timed(lambda: ctx.run_dense(pL, pR, cfg, out=pin))
    same = all(np.array_equal(o[0].view(np.uint32), want.view(np.uint32)) for o in (o1, o2, o3))

    n_pairs = 24
    one = [deir_packet(L, 1, f) + deir_packet(R, 0, f) for f in range(n_pairs)]
    st = Stream(ctx, W, H, cfg)
    st.feed(one[0])
    t0 = time.perf_counter()
    for f in range(n_pairs):
        st.feed(one[f])
    ms_stream_one = (time.perf_counter() - t0) * 1e3 / n_pairs
    blob = b"".join(one)
    st.feed(blob)
    t0 = time.perf_counter()
    outb = st.feed(blob)
    ms_stream_run = (time.perf_counter() - t0) * 1e3 / n_pairs
    pk = parse_dedd(outb)
    stream_ok = len(pk) == n_pairs and all(np.array_equal(p["q4"], pk[0]["q4"]) for p in pk)
    st.close()
    # the native fd loop (de_stream_run_fd, what `python -m doubleeye_amd.stream
    # --stream` runs) from a file to /dev/null; its set-up (three captured plans)
    # timed on an empty input and taken off
    import ctypes as C

    from doubleeye_amd._lib import check, lib

    def run_fd(path):
        fi, fo = os.open(path, os.O_RDONLY), os.open(os.devnull, os.O_WRONLY)
        pairs = C.c_long(0)
        t0 = time.perf_counter()
        check(lib().de_stream_run_fd(ctx.h, fi, fo, W, H, C.byref(cfg.to_c()), 1,
                                     C.byref(pairs), None))
        dt = time.perf_counter() - t0
        os.close(fi)
        os.close(fo)
        return dt, pairs.value
    with tempfile.TemporaryDirectory() as td:
        empty, full = os.path.join(td, "empty"), os.path.join(td, "run")
        open(empty, "wb").close()
        with open(full, "wb") as f:
            for _ in range(4):
                f.write(blob)
        run_fd(empty)
        t_set = min(run_fd(empty)[0] for _ in range(2))
        t_run, fd_pairs = run_fd(full)
    ms_fd = (t_run - t_set) * 1e3 / max(1, fd_pairs)
    for b in (dL, dR, dd, dm):
        b.free()
    ctx.close()
    out = {
        "metric": "dense 848x480 D=128 census 9x7 host-to-host ms per pair",
        "value": round(ms_pin, 4), "unit": "ms", "higher_is_better": False,
        "resident_ms": round(ms_res, 4),
        "host_pageable_fresh_ms": round(ms_fresh, 4),
        "host_pageable_reused_ms": round(ms_reuse, 4),
        "host_pinned_ms": round(ms_pin, 4),
        "dedd_stream_one_pair_a_call_ms": round(ms_stream_one, 4),
        "dedd_stream_pipelined_ms": round(ms_stream_run, 4),
        "dedd_fd_loop_ms": round(ms_fd, 4), "dedd_fd_loop_pairs": fd_pairs,
        "over_resident": {"pageable_fresh": round(ms_fresh / ms_res, 2),
                          "pageable_reused": round(ms_reuse / ms_res, 2),
                          "pinned": round(ms_pin / ms_res, 2),
                          "dedd_pipelined": round(ms_stream_run / ms_res, 2),
                          "dedd_fd_loop": round(ms_fd / ms_res, 2)},
        "outputs_equal_resident": bool(same), "dedd_packets_ok": bool(stream_ok),
        "n_gpus": 1, "steps": reps, "warmup": 3, "ms_per_step": round(ms_pin, 4),
        "scaling": "weak", "vs_baseline": None, "dtype": "i16 volume / f32 messages",
        "data": "synthetic (seeded slanted-plane pair, seed 2)",
        "config": {"workload": "dense MASDA 848x480, D=128 (d=1..128), census 9x7, "
                               "sigma_s 20, sigma_r 10/255, 8 BP iterations; host buffers"},
    }
    print(json.dumps(out), flush=True)


def bench_c2f(a):
    """The prior as a mask: one 848x480 slanted pair at D = 128 (d 0..127),
    8 iterations, through the full sweep, `--prior` (a full-resolution prior
    from the pair's ground truth, band 2) and `--c2f`, host images to host
    disparity. The CPU baseline is the reference `de_dense --c2f --threads 1`
    (oracle/_ref/de_dense) on the same pair, one run. `resident_ms` times the
    full sweep and --c2f again with the pair and outputs resident in HBM (each
    call followed by a stream wait)."""
    import numpy as np

    from doubleeye_amd import synth
    from doubleeye_amd.dense import Cfg, Context

    ctx = Context(0)
    L, R, gt = synth.slanted_pair(848, 480, seed=2)[:3]
    cfg = Cfg(dmax=127, iters=8, subpixel=0)
    prior = np.nan_to_num(gt, nan=float(np.nanmedian(gt))).astype(np.float32)
    reps = max(5, a.steps // 20)

    def timed(fn):
        fn()
        t0 = time.perf_counter()
        for _ in range(reps):
            out = fn()
        return (time.perf_counter() - t0) * 1e3 / reps, out

    ms_full, full = timed(lambda: ctx.run_dense(L, R, cfg))
    ms_prior, pri = timed(lambda: ctx.run_dense_prior(L, R, prior, cfg, 2))
    ms_c2f, c2f = timed(lambda: ctx.run_dense_c2f(L, R, cfg, 2))
    # the same with the pair and the outputs resident in HBM (no PCIe copies)
    H, W = L.shape
    dL, dR = ctx.to_device(L), ctx.to_device(R)
    dd, dm = ctx.empty((H, W), np.float32), ctx.empty((H, W), np.float32)

    def resident(fn):
        def run():
            fn()
            ctx.sync()
        return timed(run)[0]

    ms_full_dev = resident(lambda: ctx.run_dense_dev(dL.ptr, dR.ptr, W, H, cfg, dd.ptr, dm.ptr))
    ms_c2f_dev = resident(lambda: ctx.run_dense_c2f_dev(dL.ptr, dR.ptr, W, H, cfg, dd.ptr,
                                                        dm.ptr, 0, 2))
    c2f_dev_same = bool(np.array_equal(dd.numpy().view(np.uint32), c2f[0].view(np.uint32)))

    def bad1(d):
        ok = np.isfinite(d) & np.isfinite(gt)
        return round(float((np.abs(d[ok] - gt[ok]) > 1.0).mean()), 4), round(float(ok.mean()), 4)

    out = {
        "metric": "dense 848x480 D=128 host-to-host time: full sweep vs --prior vs --c2f",
        "value": round(ms_c2f, 3), "unit": "ms", "higher_is_better": False,
        "full_ms": round(ms_full, 3), "prior_ms": round(ms_prior, 3), "c2f_ms": round(ms_c2f, 3),
        "c2f_used_mask": bool(c2f[4]),
        "resident_ms": {"full": round(ms_full_dev, 3), "c2f": round(ms_c2f_dev, 3),
                        "c2f_equal_to_host_call": c2f_dev_same},
        "bad1_coverage": {"full": bad1(full[0]), "prior": bad1(pri[0]), "c2f": bad1(c2f[0])},
        "n_gpus": 1, "steps": reps, "warmup": 1, "ms_per_step": round(ms_c2f, 3),
        "scaling": "weak", "vs_baseline": None, "dtype": "int16 volume / f32 messages",
        "data": "synthetic (slanted planes 848x480, seed 2)",
        "config": {"workload": "848x480, d 0..127, 8 iterations, band 2, no sub-pixel"},
    }
    if not a.no_cpu_baseline:
        exe = os.path.join(ROOT, "oracle", "_ref", "de_dense")
        if os.path.exists(exe):
            with tempfile.TemporaryDirectory() as td:
                lp, rp, op = (os.path.join(td, n) for n in ("l.y8", "r.y8", "o.f32"))
                L.tofile(lp)
                R.tofile(rp)
                cmd = [exe, lp, rp, "848", "480", "--threads", "1", "--dmax", "127", "--iters",
                       "8", "--no-subpixel", "--c2f", "--out", op]
                t0 = time.perf_counter()
                subprocess.run(cmd, check=True, capture_output=True)
                t = time.perf_counter() - t0
            out["cpu_baseline"] = {"value": round(t * 1e3, 1), "unit": "ms", "cores": 1,
                                   "kind": "reference",
                                   "sample": "one run of de_dense --c2f --threads 1 (process "
                                             "wall time, file I/O included)"}
        else:
            out["cpu_baseline"] = {"error": "oracle/_ref/de_dense not built on this host"}
    ctx.close()
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000,
                    help="timed steps (c2: frames; 1000 frames in flight take ~0.13 s, long "
                         "enough that launch jitter averages out)")
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-budget", type=float, default=12.0)
    ap.add_argument("--workload", default="c2",
                    choices=["c2", "c3", "c4", "c5", "sparse", "dense_masda", "detect",
                             "pipe", "c2f", "full", "host"],
                    help="c2: one 848x480x128 pair (default, the headline); c4: 64-pair "
                         "batch; c3: keypoint-graph BP; c5: serving loop; sparse: keypoint "
                         "match_masda; detect: keypoint detector; pipe: de_pipe's pair; "
                         "c2f: the prior mask; full: the FP32 full-volume path (--topk 0); "
                         "host: C2 from host buffers to host buffers")
    ap.add_argument("--duration", type=float, default=60.0, help="c5: seconds of arrivals")
    ap.add_argument("--rate", type=float, default=30.0,
                    help="c5: frames per second of each of the 16 streams (Poisson)")
    ap.add_argument("--frames-in-flight", type=int, default=4, choices=[1, 2, 3, 4],
                    help="c2: frames in flight, each on its own HIP stream (1: one at a time; "
                         "4 = the hardware queues a process gets by default)")
    ap.add_argument("--census", default="both", choices=["both", "9x7"],
                    help="c2: also run the 7x7 census beside the stated 9x7 (both, the "
                         "default), or 9x7 alone (profiling runs)")
    ap.add_argument("--row-variant", type=int, default=-1,
                    help="c2, c4: force the row pass's launch variant (0-3; -1: the "
                         "library's choice; c4: the batch's plan only)")
    ap.add_argument("--col-variant", type=int, default=-1,
                    help="c2: force the column pass's launch variant (0-2: 8, 16, 32 rows in "
                         "flight; -1: the library's choice)")
    ap.add_argument("--c5-seconds", type=float, default=20.0,
                    help="c2: seconds of the 16-stream serving leg reported beside the "
                         "headline (p50/p99 latency); 0 skips it")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="first run one plain device-to-device copy of MIB MiB (the "
                         "known-size kernel that scales rocprofv3 FETCH/WRITE_SIZE)")
    a = ap.parse_args()
    if a.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path "
              "(a dense frame would shard by stream, or by disparity range with a top-2 merge)",
              file=sys.stderr)
        sys.exit(2)
    if a.workload == "c4":
        bench_c4(a)
        return
    if a.workload == "sparse":
        bench_sparse(a)
        return
    if a.workload == "c3":
        bench_c3(a)
        return
    if a.workload == "c5":
        bench_c5(a)
        return
    if a.workload == "dense_masda":
        bench_dense_masda(a)
        return
    if a.workload == "detect":
        bench_detect(a)
        return
    if a.workload == "pipe":
        bench_pipe(a)
        return
    if a.workload == "c2f":
        bench_c2f(a)
        return
    if a.workload == "full":
        bench_full(a)
        return
    if a.workload == "host":
        bench_host(a)
        return

    import numpy as np

    from doubleeye_amd import synth
    from doubleeye_amd.dense import Context

    # Device memory and streams come from the library's own HIP runtime; no
    # second runtime (e.g. torch's bundled one) is loaded into the process.
    L, R, _ = synth.slanted_pair(W, H, seed=2)
    nfl = a.frames_in_flight
    ctx = Context(0)
    # one more context per further frame in flight, for its HIP stream
    extra = [Context(0) for _ in range(nfl - 1)]
    if a.calibrate_copy:
        n = a.calibrate_copy << 20
        src = ctx.to_device(np.ones(n, np.uint8))
        dst = ctx.empty((n,), np.uint8)
        ctx.copy_dev(dst, src)          # one plain device-to-device copy of n bytes
        ctx.fill_dev(dst, 0)            # one plain device fill of n bytes
        ctx.sync()
        src.free(); dst.free()
    dL, dR = ctx.to_device(L), ctx.to_device(R)
    if a.row_variant >= 0 or a.col_variant >= 0:
        ctx.set_variants(a.row_variant, a.col_variant)

    m9 = measure_c2(ctx, extra, dL, dR, CFG, a)
    m7 = measure_c2(ctx, extra, dL, dR, CFG7, a) if a.census == "both" else None
    per = m9["stage_ms"]
    dom = max(per, key=per.get)
    ab = algorithmic_bytes(W, H, D)
    achieved = ab[dom] / (per[dom] * 1e-3) / 1e9
    frame_gbs = ab["frame"] / (m9["ms"] * 1e-3) / 1e9
    out = {
        "metric": "dense Mpx*disp/s at 848x480x128 (C2: census 9x7, 8 BP iterations)",
        "value": round(W * H * D / (m9["ms"] * 1e-3) / 1e6, 1),
        "unit": "Mpx*disp/s",
        "n_gpus": 1, "steps": a.steps, "warmup": a.warmup,
        "ms_per_step": round(m9["ms"], 4),
        "frames_per_s": round(1e3 / m9["ms"], 1),
        "single_stream_ms_per_step": round(m9["ms_single"], 4),
        "coverage": m9["coverage"],
        "deterministic": m9["deterministic"] and (m7 is None or m7["deterministic"]),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "i16 volume / f32 messages",
        "data": "synthetic (seeded slanted-plane pair, doubleeye_amd.synth.slanted_pair seed 2)",
        "config": {"workload": "dense MASDA 848x480, D=128 (d=1..128), census 9x7, "
                               "sigma_s 20, sigma_r 10/255, 8 BP iterations, top-2",
                   "W": W, "H": H, "D": D, "iters": 8, "census": "9x7",
                   "frames_in_flight": nfl, "in_flight_row_pass": m9["inflight_row_pass"]},
        "roofline": {"bound": "hbm", "kernel": f"k_{dom}", "achieved": round(achieved, 1),
                     "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": TRAFFIC.get(dom),
                     "traffic_source": TRAFFIC_SOURCE,
                     "kernel_ms": round(per[dom], 4),
                     "stage_ms": {n: round(v, 4) for n, v in per.items()},
                     "frame_algorithmic_bytes": ab["frame"],
                     "frame_achieved_gbs": round(frame_gbs, 1),
                     "frame_frac": round(frame_gbs / HBM_PEAK_GBS, 4)},
    }
    if m7 is not None:
        # the reference's own tool is fixed at census 7x7 (core/tools/de_dense.cpp:748):
        # the same workload at 7x7, timed beside the reference tool
        out["census7x7"] = {"value": round(W * H * D / (m7["ms"] * 1e-3) / 1e6, 1),
                            "ms_per_step": round(m7["ms"], 4),
                            "single_stream_ms_per_step": round(m7["ms_single"], 4),
                            "stage_ms": {n: round(v, 4) for n, v in m7["stage_ms"].items()},
                            "coverage": m7["coverage"]}
    # C1 (the reference's CPU-runnable case): 424x240, d 1..48, 5 iterations,
    # census 9x7, fronto planes at 8/20/35
    L1, R1, _ = synth.fronto_pair(424, 240, seed=1, disparities=(8, 20, 35))
    d1L, d1R = ctx.to_device(L1), ctx.to_device(R1)
    # its frames in flight take the streaming row pass (two waves a SIMD hide
    # the short rows' latencies): four in flight 0.0252 ms a frame against
    # 0.030-0.033 for the resident pass without slots (at C2: 0.135 against 0.118)
    m1 = measure_c2(ctx, extra, d1L, d1R, CFG_C1, a, W=424, H=240, inflight_row=ROW_STREAM16)
    out["c1"] = {"workload": "dense MASDA 424x240, D=48 (d=1..48), census 9x7, 5 BP iterations",
                 "value": round(424 * 240 * 48 / (m1["ms"] * 1e-3) / 1e6, 1),
                 "unit": "Mpx*disp/s", "ms_per_step": round(m1["ms"], 4),
                 "single_stream_ms_per_step": round(m1["ms_single"], 4),
                 "stage_ms": {n: round(v, 4) for n, v in m1["stage_ms"].items()},
                 "coverage": m1["coverage"], "deterministic": m1["deterministic"],
                 "in_flight_row_pass": m1["inflight_row_pass"]}
    d1L.free(); d1R.free()
    for c in extra:
        c.close()
    ctx.close()
    if not a.no_cpu_baseline:
        out["c1"]["cpu_baseline"] = cpu_baseline_port(np.ascontiguousarray(L1),
                                                       np.ascontiguousarray(R1),
                                                       min(3.0, a.cpu_budget), CFG_C1,
                                                       w=424, h=240)
        Lc, Rc = np.ascontiguousarray(L), np.ascontiguousarray(R)
        out["cpu_baseline"] = cpu_baseline_port(Lc, Rc, a.cpu_budget)
        if m7 is not None:
            out["census7x7"]["cpu_reference"] = cpu_baseline(Lc, Rc, a.cpu_budget / 2)
    if a.c5_seconds > 0:
        # the second half of the metric: p50/p99 frame latency at 16 mixed streams
        c5 = run_c5(a.c5_seconds)
        out["c5"] = {k: c5[k] for k in ("p50_ms", "p99_ms", "max_ms", "frames_per_s",
                                         "dense_mpx_disp_per_s", "mean_batch", "max_batch",
                                         "frames", "dropped_busy",
                                         "sampled_results_identical")}
        out["c5"]["seconds"] = a.c5_seconds
        out["p50_ms"] = c5["p50_ms"]
        out["p99_ms"] = c5["p99_ms"]
        out["mean_batch"] = c5["mean_batch"]
        # the same 16 streams offered 400 Hz each (6400 frames/s): the server
        # falls behind, so frames queue and leave in batches
        sat = run_c5(min(5.0, a.c5_seconds), rate_hz=400.0)
        out["c5_saturated"] = {k: sat[k] for k in ("p50_ms", "p99_ms", "frames_per_s",
                                                    "offered_frames_per_s",
                                                    "dense_mpx_disp_per_s", "mean_batch",
                                                    "max_batch", "frames", "dropped_busy",
                                                    "sampled_results_identical")}
        out["c5_saturated"]["seconds"] = min(5.0, a.c5_seconds)
    print(json.dumps(out), flush=True)


def measure_c2(ctx, extra, dL, dR, cfg_kw: dict, a, W: int = W, H: int = H,
               inflight_row: int = ROW_NO_SLOTS) -> dict:
    """One configuration of a W x H pair (C2 by default): frames in flight (a
    captured plan per frame slot, slot k on HIP stream k), then one frame at a
    time, then the per-stage kernel times from hipEvents between the launches
    on the launching stream (one frame at a time, so each kernel runs alone)."""
    import numpy as np

    from doubleeye_amd.dense import Cfg, Pair

    nfl = len(extra) + 1
    cfg = Cfg(**cfg_kw)
    disp = ctx.empty((H, W), np.float32)
    margin = ctx.empty((H, W), np.float32)

    def step():
        ctx.run_dense_dev(dL.ptr, dR.ptr, W, H, cfg, disp.ptr, margin.ptr)

    step()
    ctx.sync()
    first = disp.numpy()
    slots = []
    # Frames in flight take the register-resident row pass without LDS
    # forward slots (its workgroups then hold little LDS, so other frames'
    # kernels fit beside them); one frame alone takes the library's choice.
    # Measured, ms a frame with 2 / 3 / 4 in flight at 848x480x128: resident
    # with LDS slots 0.123 / 0.120 / 0.119, without 0.118 / 0.122 / 0.119,
    # streaming (about 210 registers a lane, two waves a SIMD) 0.135 / 0.127 /
    # 0.123; at 424x240x48 with four in flight 0.0358 / 0.0281 / 0.0277.
    inflight_rv = a.row_variant if a.row_variant >= 0 else (inflight_row if nfl > 1 else -1)
    ctx.set_variants(inflight_rv, a.col_variant)
    for _ in range(nfl):
        o = (ctx.empty((H, W), np.float32), ctx.empty((H, W), np.float32))
        slots.append((ctx.plan([Pair(dL.ptr, dR.ptr, W, H, o[0].ptr, o[1].ptr)], [cfg]), o))
    ctx.set_variants(a.row_variant, a.col_variant)
    streams = [ctx.stream] + [c.stream for c in extra]

    def run_inflight(k_steps):
        for k in range(k_steps):
            slots[k % nfl][0].launch(streams[k % nfl])

    def sync_all():
        ctx.sync()
        for c in extra:
            c.sync()

    run_inflight(a.warmup)
    sync_all()
    t0 = time.perf_counter()
    run_inflight(a.steps)
    sync_all()
    ms = (time.perf_counter() - t0) * 1e3 / a.steps
    deterministic = all(bool(np.array_equal(first.view(np.uint32), o[0].numpy().view(np.uint32)))
                        for _, o in slots)
    last = slots[0][1][0].numpy()
    for _ in range(a.warmup):
        step()
    ctx.sync()
    t0 = time.perf_counter()
    for _ in range(a.steps):
        step()
    ctx.sync()
    ms_single = (time.perf_counter() - t0) * 1e3 / a.steps
    deterministic = deterministic and bool(
        np.array_equal(first.view(np.uint32), disp.numpy().view(np.uint32)))
    ctx.set_profiling(True)
    for _ in range(max(20, a.steps // 4)):
        step()
    ctx.sync()
    stage_ms, runs = ctx.stage_ms()
    ctx.set_profiling(False)
    names = ["prep", "row", "col", "solve"]
    per = {n: stage_ms[i] / max(1, runs) for i, n in enumerate(names)}
    for p, o in slots:
        p.close()
        o[0].free(); o[1].free()
    disp.free(); margin.free()
    return {"ms": ms, "ms_single": ms_single, "stage_ms": per, "deterministic": deterministic,
            "inflight_row_pass": {-1: "library's choice", 0: "resident",
                                  1: "resident, no LDS slots", 2: "streaming, prefetch 8",
                                  3: "streaming, prefetch 16"}[inflight_rv],
            "coverage": round(float(np.isfinite(last).mean()), 4)}


if __name__ == "__main__":
    main()
