import os
import sys

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
