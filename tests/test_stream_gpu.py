"""DEIR -> DEDD stream front end (de_stream_*, python -m doubleeye_amd.stream):
pairing by frame number as de_dense_cuda --stream does, and every DEDD packet
equal to the plain single-pair result (Q4 disparity and u8 confidence)."""
import os
import subprocess
import sys

import numpy as np
import pytest

from conftest import ROOT, check_conf_bytes

pytestmark = pytest.mark.gpu
W, H = 200, 96
KW = dict(dmax=48, iters=3, sigma_s=20.0, sigma_r=0.0392)


def note(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def plain(ctx, L, R, cfg):
    from doubleeye_amd.dense import Pair
    dL, dR = ctx.to_device(L), ctx.to_device(R)
    disp = ctx.empty((H, W), np.float32)
    margin = ctx.empty((H, W), np.float32)
    conf = ctx.empty((H, W), np.uint8)
    q4 = ctx.empty((H, W), np.int16)
    ctx.dense_batch_dev([Pair(dL.ptr, dR.ptr, W, H, disp.ptr, margin.ptr, 0, conf.ptr, q4.ptr)],
                        [cfg])
    ctx.sync()
    return q4.numpy(), conf.numpy()


def packets_and_model():
    """A DEIR byte stream with out-of-order halves, a replaced image, and more
    than 8 lone frames; plus the pairs a de_dense_cuda-style pairer emits."""
    from doubleeye_amd import synth
    from doubleeye_amd.stream import deir_packet
    frames = {n: synth.fronto_pair(W, H, 20 + n)[:2] for n in range(6)}
    seq = []
    seq.append((frames[0][0], 1, 0)); seq.append((frames[0][1], 0, 0))
    seq.append((frames[1][1], 0, 1)); seq.append((frames[1][0], 1, 1))        # right first
    junk = np.full((H, W), 7, np.uint8)
    seq.append((junk, 1, 2)); seq.append((frames[2][0], 1, 2))               # replaced left
    seq.append((frames[2][1], 0, 2))
    for n in range(100, 110):                                                # lone frames
        seq.append((junk, 1, n))
    seq.append((frames[3][0], 1, 3)); seq.append((frames[3][1], 0, 3))
    seq.append((frames[4][1], 0, 4)); seq.append((frames[4][0], 1, 4))
    data = b"".join(deir_packet(img, st, fr) for img, st, fr in seq)
    pending, out, dropped = {}, [], 0
    for img, st, fr in seq:
        if fr not in pending:
            pending[fr] = (img, st)
            while len(pending) > 8:
                del pending[min(pending)]
                dropped += 1
            continue
        if pending[fr][1] == st:
            pending[fr] = (img, st)
            continue
        left = img if st == 1 else pending[fr][0]
        right = pending[fr][0] if st == 1 else img
        out.append((fr, left, right))
        del pending[fr]
    return data, out, dropped


def check_packets(ctx, got, want, oracle=None):
    """Every packet equals the plain call; with the oracle, its Q4 and
    confidence planes also equal the oracle's DEDD planes (de_dense_cuda.cu:
    731-745) away from exact-tie pixels, confidence bytes to the .5-boundary
    rule."""
    from doubleeye_amd.dense import Cfg
    assert [p["frame"] for p in got] == [w[0] for w in want]
    for p, (fr, L, R) in zip(got, want):
        q4, conf = plain(ctx, L, R, Cfg(**KW))
        assert np.array_equal(p["left"], L), fr
        assert np.array_equal(p["q4"], q4), fr
        assert p["conf"] is not None and np.array_equal(p["conf"], conf), fr
        if oracle is not None:
            od, _, oc, ties = oracle.dense_run(L, R, oracle.cfg(**KW), want_cand=True,
                                               want_ties=True)
            oq4, oconf, p255 = oracle.dedd_planes(od, oc, -0.1)
            keep = ~ties.astype(bool)
            assert np.array_equal(p["q4"][keep], oq4[keep]), fr
            check_conf_bytes(np.where(keep, p["conf"], 0), np.where(keep, oconf, 0),
                             np.where(keep, p255, 0), (p["q4"] != -32768) & keep,
                             f"DEDD frame {fr}")


def test_feed_in_chunks(ctx, oracle):
    from doubleeye_amd.dense import Cfg
    from doubleeye_amd.stream import Stream, parse_dedd
    data, want, dropped = packets_and_model()
    s = Stream(ctx, W, H, Cfg(**KW))
    out = b""
    rng = np.random.default_rng(3)
    pos = 0
    while pos < len(data):
        n = int(rng.integers(1, 40000))
        out += s.feed(data[pos:pos + n])
        pos += n
    note(f"[stream] {len(want)} pairs")
    check_packets(ctx, parse_dedd(out), want, oracle)
    assert s.stats() == (len(want), dropped)
    s.close()


def test_feed_whole_stream_at_once(ctx):
    """Every pair of the stream in one feed call: more pairs than slots in
    flight, so slots are reused while earlier pairs are still on the GPU."""
    from doubleeye_amd.dense import Cfg
    from doubleeye_amd.stream import Stream, parse_dedd
    data, want, dropped = packets_and_model()
    s = Stream(ctx, W, H, Cfg(**KW))
    out = s.feed(data)
    note(f"[stream] one feed, {len(want)} pairs")
    check_packets(ctx, parse_dedd(out), want)
    assert s.stats() == (len(want), dropped)
    s.close()


def test_bad_magic_raises(ctx):
    from doubleeye_amd._lib import DeError
    from doubleeye_amd.dense import Cfg
    from doubleeye_amd.stream import Stream
    s = Stream(ctx, W, H, Cfg(**KW))
    with pytest.raises(DeError):
        s.feed(b"XXXX" + bytes(40))
    s.close()


def test_command_line_stream_and_file_modes(ctx, tmp_path):
    from doubleeye_amd.dense import Cfg
    from doubleeye_amd.stream import parse_dedd
    data, want, _ = packets_and_model()
    args = [sys.executable, "-m", "doubleeye_amd.stream", "/dev/null", "/dev/null", str(W), str(H),
            "--stream", "--threads", "4", "--dmax", "48", "--iters", "3", "--sigma-s", "20",
            "--sigma-r", "0.0392"]
    note("[stream] command line")
    r = subprocess.run(args, input=data, capture_output=True, cwd=ROOT, timeout=120)
    assert r.returncode == 0, r.stderr.decode()
    check_packets(ctx, parse_dedd(r.stdout), want)
    fr, L, R = want[0]
    lp, rp, op = (str(tmp_path / n) for n in ("l.y8", "r.y8", "d.f32"))
    L.tofile(lp)
    R.tofile(rp)
    r = subprocess.run([sys.executable, "-m", "doubleeye_amd.stream", lp, rp, str(W), str(H),
                        "--dmax", "48", "--iters", "3", "--sigma-s", "20", "--sigma-r", "0.0392",
                        "--out", op], capture_output=True, cwd=ROOT, timeout=120)
    assert r.returncode == 0, r.stderr.decode()
    d, _, _ = ctx.run_dense(L, R, Cfg(**KW))
    got = np.fromfile(op, np.float32).reshape(H, W)
    assert np.array_equal(got.view(np.uint32), d.view(np.uint32)) or \
        np.array_equal(np.isnan(got), np.isnan(d)) and np.array_equal(got[~np.isnan(got)], d[~np.isnan(d)])
