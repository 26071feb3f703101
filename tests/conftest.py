import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)
GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")


@pytest.fixture(scope="session")
def oracle():
    from oracle import binding
    binding.lib()
    return binding


@pytest.fixture(scope="session")
def ctx():
    from doubleeye_amd.dense import Context
    c = Context(0)
    yield c
    c.close()


def flags_to_kwargs(flags: str) -> dict:
    """de_dense command-line flags -> config keyword arguments."""
    t = flags.split()
    kw, i = {}, 0
    while i < len(t):
        a = t[i]
        if a == "--dmax": kw["dmax"] = int(t[i + 1]); i += 2
        elif a == "--iters": kw["iters"] = int(t[i + 1]); i += 2
        elif a == "--sigma-s": kw["sigma_s"] = float(t[i + 1]); i += 2
        elif a == "--sigma-r": kw["sigma_r"] = float(t[i + 1]); i += 2
        elif a == "--min-margin": kw["min_margin"] = float(t[i + 1]); i += 2
        elif a == "--min-ratio": kw["min_ratio"] = float(t[i + 1]); i += 2
        elif a == "--no-subpixel": kw["subpixel"] = 0; i += 1
        elif a == "--fit-parabola": kw["fit_eq"] = 0; i += 1
        elif a == "--lrc": kw["lrc"] = 1; i += 1
        elif a == "--band": kw["band"] = int(t[i + 1]); i += 2
        elif a == "--c2f": kw["c2f"] = 1; i += 1
        else: raise ValueError(a)
    return kw


CONF_BOUNDARY = 2e-3   # |P*255 - (k + 0.5)| where a confidence byte may round the other way
CONF_MAX_FRAC = 1e-3   # differing confidence bytes per answered pixel


def check_conf_bytes(conf, oconf, p255, answered, tag):
    """u8 confidence bytes against the oracle's: the device's logf/expf and
    glibc's differ by about an ulp on some arguments, so a byte may round the
    other way where P*255 lies next to a .5 boundary, and nowhere else.
    Returns the number of differing bytes."""
    diff = conf.astype(np.int32) - oconf.astype(np.int32)
    bad = diff != 0
    n = int(bad.sum())
    dist = np.abs(p255 - (np.floor(p255) + np.float32(0.5)))
    worst = float(dist[bad].max()) if n else 0.0
    print(f"[conf] {tag}: {n} of {int(answered.sum())} answered bytes differ by one "
          f"(largest distance of P*255 from k+0.5 among them {worst:.2e})", flush=True)
    assert (np.abs(diff) <= 1).all(), f"{tag}: a byte differs by more than one"
    assert (dist[bad] <= CONF_BOUNDARY).all(), f"{tag}: a byte differs away from a .5 boundary"
    assert n <= max(1, CONF_MAX_FRAC * answered.sum()), f"{tag}: {n} differing bytes"
    return n
