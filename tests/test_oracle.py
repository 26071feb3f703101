"""Pins the CPU oracle (oracle/cpu_ref.cpp) against the reference's own outputs
and known answers, and against closed forms. CPU only.

Sources of truth:
  * tests/golden/scenes.npz   -- the reference's six IR scene pairs and their
    meta.txt coverage_pct known answers (article/scenes/*/meta.txt), plus sha256
    of `de_dense --dmax 64 --threads 1 --out`.
  * tests/golden/dense_c1.npz -- `de_dense --threads 1` outputs on a seeded
    424x240 pair for six flag sets.
  * tests/golden/volume_f32.npz -- `de_dense --dump-vol` (FP32 volume).
  * core/tests/test_confidence.cpp:52-56, 117-129 golden values.
  * core/tests/test_preproc.cpp:50-93 census invariants.
"""
import hashlib
import os

import numpy as np
import pytest

from conftest import GOLDEN, flags_to_kwargs

SCENES = ["night_on", "night_off", "evening_on", "evening_off", "morning_on", "morning_off"]


@pytest.fixture(scope="module")
def scenes():
    return np.load(os.path.join(GOLDEN, "scenes.npz"))


@pytest.fixture(scope="module")
def c1():
    return np.load(os.path.join(GOLDEN, "dense_c1.npz"))


def sha(a: np.ndarray) -> str:
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


@pytest.mark.parametrize("scene", SCENES)
def test_scene_known_answer(oracle, scenes, scene):
    L, R = scenes[f"{scene}_L"], scenes[f"{scene}_R"]
    d, _, _, _ = oracle.dense_run(L, R, oracle.cfg(dmax=64))
    cov = round(100.0 * np.isfinite(d).mean(), 1)
    assert cov == float(scenes[f"{scene}_coverage_pct"])
    assert sha(d) == str(scenes[f"{scene}_disp_sha256"]), "differs from de_dense --threads 1"


@pytest.mark.parametrize("name", ["c1", "default48", "lrc_it2", "nosub_margin",
                                  "parabola_ratio", "d96_it8"])
def test_c1_flags_bit_exact(oracle, c1, name):
    kw = flags_to_kwargs(str(c1[f"{name}_flags"]))
    d, _, cand, _ = oracle.dense_run(c1["L"], c1["R"], oracle.cfg(**kw), want_cand=True)
    assert sha(d) == str(c1[f"{name}_disp_sha256"])
    assert sha(cand) == str(c1[f"{name}_cand_sha256"])
    if name == "c1":
        np.testing.assert_array_equal(d.view(np.uint32), c1["c1_disp"].view(np.uint32))


def test_volume_f32_matches_dump(oracle):
    z = np.load(os.path.join(GOLDEN, "volume_f32.npz"))
    kw = flags_to_kwargs(str(z["flags"]))
    vol = oracle.volume_f32(z["L"], z["R"], oracle.cfg(**kw))
    ref = z["vol"]
    # the float filter is restated in the reference's operation order with no
    # FMA contraction, so the volume is equal bit for bit
    assert np.array_equal(vol.view(np.uint32), ref.view(np.uint32))


FULL = np.load(os.path.join(GOLDEN, "dense_full.npz"))
FULL_CASES = [(p, n) for p in ("fronto", "slanted") for n in ("it4", "it0", "margin", "nosub")]


@pytest.mark.parametrize("case", FULL_CASES, ids=["-".join(c) for c in FULL_CASES])
def test_full_volume_solver_matches_reference(case, oracle):
    """`de_dense --topk 0` (solve_row, de_dense.cpp:566-712): disparities equal
    except at pixels whose claim ties at an exactly equal belief."""
    p, n = case
    kw = flags_to_kwargs(str(FULL[f"{p}_{n}_flags"]).replace("--topk 0", ""))
    d, _, ties = oracle.dense_run_full(FULL[f"{p}_L"], FULL[f"{p}_R"], oracle.cfg(**kw),
                                       want_ties=True)
    ref = FULL[f"{p}_{n}_disp"]
    same = (d.view(np.uint32) == ref.view(np.uint32)) | (np.isnan(d) & np.isnan(ref))
    assert (same | (ties > 0)).all()


def test_census_bits_and_invariants(oracle):
    rng = np.random.default_rng(7)
    img = rng.integers(0, 256, (40, 53), dtype=np.uint8)
    # bit count: on a strictly increasing ramp every neighbour before the
    # centre in scan order is smaller -> exactly the low nbits/2 bits are set
    ramp = np.arange(40 * 53, dtype=np.int64).reshape(40, 53)
    ramp = (ramp * 255 // ramp.max()).astype(np.uint8)
    for hw, hh, nbits in ((3, 3, 48), (4, 3, 62)):
        c = oracle.census(img, hw, hh)
        assert int(c.max()) < (1 << nbits)
        cr = oracle.census(np.arange(9 * 9, dtype=np.uint8).reshape(9, 9), hw, hh)
        assert int(cr[4, 4]) == (1 << (nbits // 2)) - 1
        # uniform image -> all zero (test_preproc.cpp:59-63)
        assert not oracle.census(np.full((20, 20), 77, np.uint8), hw, hh).any()
        # monotonic invariance (test_preproc.cpp:68-93): a strictly increasing
        # intensity map leaves every descriptor unchanged
        small = (img // 2).astype(np.uint8)                      # 0..127
        lut = np.sort(np.random.default_rng(hw).choice(256, 128, replace=False)).astype(np.uint8)
        assert np.array_equal(oracle.census(small, hw, hh), oracle.census(lut[small], hw, hh))
        # brute force at a few pixels
        for (y, x) in ((hh, hw), (20, 25), (40 - hh - 1, 53 - hw - 1)):
            v, b = 0, 0
            for dy in range(-hh, hh + 1):
                for dx in range(-hw, hw + 1):
                    if dx == 0 and dy == 0:
                        continue
                    if img[y + dy, x + dx] < img[y, x]:
                        v |= 1 << b
                    b += 1
            assert int(c[y, x]) == v
        # border is zero
        assert not c[:hh].any() and not c[-hh:].any()
        assert not c[:, :hw].any() and not c[:, -hw:].any()
    assert ramp.max() == 255


def test_confidence_goldens(oracle):
    lam = -0.1
    for s1, s2, want in ((0.9, 0.3, 1.000000), (0.5, 0.49, 0.700536),
                         (0.682, 0.537, 0.975261), (0.2, -0.5, 0.961586),
                         (0.3, -1e30, 0.986346)):
        assert abs(oracle.confidence(s1, s2, lam) - want) <= 1e-5
    assert oracle.confidence(-1e30, -1e30, lam) == 0.0
    assert oracle.confidence_u8(0.9, 0.3, lam) == 255
    assert oracle.confidence_u8(0.5, 0.49, lam) == 179
    assert abs(oracle.confidence(0.9, 0.3, lam, 9.0) - 0.35) <= 1e-6
    assert oracle.confidence(0.9, 0.3, lam, float("nan")) == oracle.confidence(0.9, 0.3, lam, 0.0)


def test_filter_step_offset_identity():
    """The device form of one recursive-filter step (offset u16, two
    multiply-adds, high-half read) equals the reference step
    f = c + ((a*(f-c) + 2^14) >> 15) (core/tools/de_dense.cpp:411-430)."""
    rng = np.random.default_rng(0)
    n = 3_000_000
    a = np.concatenate([rng.integers(0, 32768, n), [0, 0, 32767, 32767, 1, 16384]])
    c = np.concatenate([rng.integers(-16384, 16385, n), [-16384, 16384] * 3])
    f = np.concatenate([rng.integers(-16384, 16385, n), [16384, -16384] * 3])
    ref = c + ((a * (f - c) + (1 << 14)) >> 15)
    cp, fp = c + 16384, f + 16384
    acc = (65536 - 2 * a) * cp + 32768          # v_mad_u32_u24
    t2 = (2 * a) * fp + acc                      # v_mad_u32_u16, op_sel hi
    assert t2.max() < 2 ** 32 and acc.max() < 2 ** 32
    assert np.array_equal((t2 >> 16) - 16384, ref)
    assert ref.min() >= -16384 and ref.max() <= 16384
    # the row pass's form: one filter word wz = 2a << 16 | min(65536 - 2a, 65535)
    # and two 16 x 16 + 32 multiply-adds; at a = 0 the capped weight 65535 gives
    # the same high half (65535 c' + 32768 >> 16 = c' for c' <= 32768)
    wz = ((2 * a) << 16) | np.minimum(65536 - 2 * a, 65535)
    acc16 = (wz & 0xFFFF) * cp + 32768
    t2w = (wz >> 16) * fp + acc16
    assert t2w.max() < 2 ** 32
    assert np.array_equal((t2w >> 16) - 16384, ref)
    zero = a == 0
    assert zero.sum() >= 2 and np.array_equal((t2w[zero] >> 16) - 16384, c[zero])


def _seq_top2(vals, kmax):
    b0 = b1 = -32768
    i0 = i1 = -1
    for k in range(kmax + 1):
        v = vals[k]
        if v <= b1:
            continue
        if v > b0:
            b1, i1, b0, i0 = b0, i0, v, k
        else:
            b1, i1 = v, k
    return b0, i0, b1, i1


def test_packed_top2_butterfly_semantics():
    """The column pass's wave64 reduction: lane-local insert over G groups, then
    a 6-round xor butterfly on packed keys ((v'+1) << 8 | 255-k). Simulated
    here against the sequential ascending-k strictly-greater scan that defines
    the semantics (core/tools/de_dense.cpp:1286-1293), tie-saturated and spread,
    like core/tests/test_top2_merge.cpp."""
    rng = np.random.default_rng(20260809)
    for spread in (3, 12, 30000):
        for _ in range(400):
            G = int(rng.integers(1, 5))
            D = min(256, G * 64 - int(rng.integers(0, 5)))
            v = rng.integers(0, spread, 256) - spread // 2
            v = np.clip(v, -16384, 16384)
            kmax = int(rng.integers(-1, D))
            keys = np.zeros((64, G), np.int64)
            for g in range(G):
                kk = g * 64 + np.arange(64)
                ok = kk <= kmax
                keys[ok, g] = ((v[kk[ok]] + 16384 + 1) << 8) | (255 - kk[ok])
            p0 = np.zeros(64, np.int64)
            p1 = np.zeros(64, np.int64)
            for g in range(G):
                k = keys[:, g]
                p1 = np.maximum(p1, np.minimum(p0, k))
                p0 = np.maximum(p0, k)
            lanes = np.arange(64)
            for off in (1, 2, 4, 8, 16, 32):
                o0, o1 = p0[lanes ^ off], p1[lanes ^ off]
                n1 = np.maximum(np.minimum(p0, o0), np.maximum(p1, o1))
                p0 = np.maximum(p0, o0)
                p1 = n1
            assert (p0 == p0[0]).all() and (p1 == p1[0]).all()
            b0, i0, b1, i1 = _seq_top2(v, kmax)
            got = (((p0[0] >> 8) - 1 - 16384) if p0[0] else -32768,
                   (255 - (p0[0] & 255)) if p0[0] else -1,
                   ((p1[0] >> 8) - 1 - 16384) if p1[0] else -32768,
                   (255 - (p1[0] & 255)) if p1[0] else -1)
            assert got == (b0, i0, b1, i1)


def _brute_census(img, y, x, hw, hh):
    """Census bits by enumeration: bit b in scan order dy-major, dx-minor,
    centre skipped, set when the neighbour is darker than the centre."""
    v, b = 0, 0
    for dy in range(-hh, hh + 1):
        for dx in range(-hw, hw + 1):
            if dx == 0 and dy == 0:
                continue
            if img[y + dy, x + dx] < img[y, x]:
                v |= 1 << b
            b += 1
    return v, b


def test_score_table_9x7_closed_forms(oracle):
    """The 9x7 window has 62 bits, so the census level is (31 - h) / 31 in Q14
    truncated toward zero: +1 at h = 0, 0 at h = 31, -1 at h = 62, strictly
    decreasing, antisymmetric about h = 31, within one unit of the exact
    rational. The reference hard-codes the 48-bit table (half = 24,
    de_dense.cpp:1201-1210); a table built with the wrong half fails here."""
    nbits = 9 * 7 - 1
    half = nbits // 2
    assert half == 31
    tbl, adt, wq = oracle.score_tables(half, 0.15, 10)
    t = tbl[:nbits + 1].astype(np.int64)
    assert t[0] == 16384 and t[half] == 0 and t[nbits] == -16384
    assert (np.diff(t) < 0).all()
    assert np.array_equal(t, -t[::-1])
    exact = (half - np.arange(nbits + 1)) * 16384 / half
    assert (np.abs(t - exact) < 1).all() and (np.abs(t) <= np.abs(exact)).all()
    # the 7x7 table has the same shape at half = 24
    t7, _, _ = oracle.score_tables(24, 0.15, 10)
    assert t7[0] == 16384 and t7[24] == 0 and t7[48] == -16384
    # truncated-AD level: +1 at |dI| = 0, -1 from T on; blend weight int(0.15 * 1024)
    assert adt[0] == 16384 and (adt[10:] == -16384).all() and (np.diff(adt[:11]) < 0).all()
    assert wq == 153


def test_graded_score_9x7_brute_force(oracle):
    """Raw graded scores of the 9x7 path (a zero filter coefficient, sigma_s ->
    0, leaves every plane unfiltered) against a brute-force census, Hamming
    distance and blend at every pixel, and the valid window y in [3, H-3),
    x in [4 + d, W - 4) with zero outside (de_dense.cpp:1167-1243 with the
    window's own border)."""
    from doubleeye_amd import synth
    L, R, _ = synth.fronto_pair(60, 24, seed=3, disparities=(2, 5, 9))
    H, W = L.shape
    hw, hh, dmin, dmax = 4, 3, 1, 12
    assert not oracle.rf_table(0.01, 0.0392).any()      # a = 0: f = c at every step
    planes = oracle.filtered_planes(L, R, oracle.cfg(dmin=dmin, dmax=dmax, sigma_s=0.01,
                                                     census_hw=hw, census_hh=hh), 0, dmax)
    half, wq, T = 31, 153, 10
    cl = {}
    cr = {}
    for y in range(hh, H - hh):
        for x in range(hw, W - hw):
            cl[y, x], nb = _brute_census(L.astype(int), y, x, hw, hh)
            cr[y, x], _ = _brute_census(R.astype(int), y, x, hw, hh)
    assert nb == 62
    for k in range(dmax):
        d = dmin + k
        want = np.zeros((H, W), np.int64)
        for y in range(hh, H - hh):
            for x in range(hw + d, W - hw):
                h = bin(cl[y, x] ^ cr[y, x - d]).count("1")
                num = (half - h) * 16384
                c = abs(num) // half * (1 if num >= 0 else -1)      # C++ truncation
                a = 16384 - (2 * 16384 * min(abs(int(L[y, x]) - int(R[y, x - d])), T)) // T
                want[y, x] = (c * (1024 - wq) + a * wq) >> 10
        np.testing.assert_array_equal(planes[k], want, err_msg=f"plane d={d}")
    assert (planes[:, :, :hw + dmin] == 0).all() and (planes[:, :, W - hw:] == 0).all()


def test_dedd_planes_oracle(oracle):
    """Q4 = lround(16 d) for d > 0 finite, -32768 otherwise; confidence 0 where
    Q4 is none, else confidence_u8 with NaN lrc read as 0
    (de_dense_cuda.cu:731-745); goldens of test_confidence.cpp:52-56, 117-119."""
    disp = np.array([1.03125, 2.0, -1.0, np.nan, 0.0, 7.96875, 3.5], np.float32)
    c6 = np.full((6, disp.size), -1e30, np.float32)
    c6[0] = [0.9, 0.5, 0.9, 0.9, 0.9, 0.682, 0.9]
    c6[1] = [0.3, 0.49, 0.3, 0.3, 0.3, 0.537, 0.3]
    c6[5] = [np.nan, 0.0, 0.0, 0.0, 0.0, np.nan, 9.0]
    q4, conf, p255 = oracle.dedd_planes(disp, c6, -0.1)
    assert q4.tolist() == [17, 32, -32768, -32768, -32768, 128, 56]
    assert conf.tolist()[:5] == [255, 179, 0, 0, 0]
    assert conf[5] == round(0.975261 * 255) and conf[6] == round(0.35 * 255)
    assert p255[2] == 0 and abs(p255[1] - 0.700536 * 255) < 1e-3
