"""The C-ABI library loads and exports every function include/doubleeye_hip.h
declares; the ctypes struct matches the C layout. No GPU calls."""
import ctypes as C
import os
import re

from conftest import ROOT


def declared_functions():
    names = set()
    for fn in os.listdir(os.path.join(ROOT, "include")):
        if not fn.endswith(".h"):
            continue
        txt = open(os.path.join(ROOT, "include", fn)).read()
        txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
        for m in re.finditer(r"^\s*(?:const\s+)?[\w\s\*]+?\b(de_\w+)\s*\(", txt, flags=re.M):
            names.add(m.group(1))
    return names


def test_header_declares_entry_points():
    names = declared_functions()
    for must in ("de_ctx_create", "de_dense_run", "de_dense_run_dev", "de_census"):
        assert must in names


def test_library_exports_every_declared_symbol():
    from doubleeye_amd import _lib
    L = _lib.lib()
    missing = [n for n in sorted(declared_functions()) if not hasattr(L, n)]
    assert not missing, f"declared but not exported: {missing}"


def test_cfg_struct_layout_and_defaults():
    from doubleeye_amd import _lib
    from doubleeye_amd.dense import Cfg
    assert C.sizeof(_lib.DeCfgC) == 17 * 4
    c = _lib.DeCfgC()
    _lib.lib().de_default_cfg(C.byref(c))     # pure host function, no GPU
    py = Cfg().to_c()
    for name, _ in _lib.DeCfgC._fields_:
        assert getattr(c, name) == getattr(py, name), name
