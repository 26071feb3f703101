"""Builds doubleeye_amd/libdoubleeye_hip.so for gfx950 with hipcc (in-tree).

    python -m doubleeye_amd.build

Each .hip file is compiled to an object of its own (the row solver with
-ffp-contract=off, so its float arithmetic rounds once per operation like the
reference's x86 build), then linked into one shared library.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libdoubleeye_hip.so")
ARCH = os.environ.get("DE_OFFLOAD_ARCH", "gfx950")

COMMON = [
    f"--offload-arch={ARCH}", "-O3", "-fPIC", "-std=c++17", "-Wall",
    # no packed fp32 instructions anywhere: some of their forms have given wrong
    # results on this hardware with no error
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]
SOURCES = {
    "de_dense.hip": [],
    "de_row.hip": [],
    "de_col.hip": [],
    "de_solve.hip": ["-ffp-contract=off"],
    "de_api.hip": ["-ffp-contract=off"],
    "de_sparse.hip": ["-ffp-contract=off"],
    "de_knn.hip": ["-ffp-contract=off"],
    "de_serve.hip": [],
    "de_edges.hip": ["-ffp-contract=off"],
    "de_fp32.hip": ["-ffp-contract=off"],
    "de_stream.hip": [],
    "de_hostio.hip": [],
    "de_detect.hip": ["-ffp-contract=off"],
    "de_prior.hip": [],
    "de_noalloc.hip": [],
}


COMPILE_TIMEOUT_S = 900   # per source file; hipcc takes seconds on each


def hipcc() -> str:
    for p in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if p and os.path.exists(p):
            return p
    raise RuntimeError("hipcc not found")


def build(verbose: bool = False) -> str:
    os.makedirs(BUILD, exist_ok=True)
    cc = hipcc()
    procs = []
    objs = []
    for src, extra in SOURCES.items():
        obj = os.path.join(BUILD, src.replace(".hip", ".o"))
        objs.append(obj)
        cmd = [cc, *COMMON, *extra, "-c", os.path.join(CSRC, src), "-o", obj]
        if verbose:
            print(" ".join(cmd))
        procs.append((src, subprocess.Popen(cmd, stdout=subprocess.PIPE,
                                            stderr=subprocess.STDOUT, text=True)))
    failed = []
    for src, p in procs:
        try:
            out, _ = p.communicate(timeout=COMPILE_TIMEOUT_S)
        except subprocess.TimeoutExpired:
            p.kill()
            out, _ = p.communicate()
            out += f"\nhipcc exceeded {COMPILE_TIMEOUT_S} s on {src}"
            p.returncode = -1
        # hipcc's host pass does not know the gfx feature flag: harmless
        lines = [ln for ln in out.splitlines() if "packed-fp32-ops" not in ln
                 and "is not a recognized feature" not in ln]
        if p.returncode != 0:
            failed.append(src)
            sys.stderr.write("\n".join(lines) + "\n")
        elif verbose and lines:
            print("\n".join(lines))
    if failed:
        raise RuntimeError(f"hipcc failed for {failed}")
    tmp = LIB + ".tmp"
    cmd = [cc, f"--offload-arch={ARCH}", "-shared", "-fPIC", *objs, "-o", tmp]
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
