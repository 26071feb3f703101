"""MI355X-native (gfx950 HIP) DoubleEye stereo-matching hot path.

The compute lives in libdoubleeye_hip.so (C ABI, include/doubleeye_hip.h);
this package is the host-side mirror of the reference's entry points:

* ``doubleeye_amd.dense``  -- run_dense / Cfg / census (core/tools/de_dense.cpp)
* ``doubleeye_amd.synth``  -- seeded synthetic stereo pairs
"""
from ._lib import DeError  # noqa: F401

__all__ = ["DeError"]
