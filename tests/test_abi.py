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


def test_deir_dedd_packet_helpers():
    """Packet helpers against hand-built bytes (no GPU)."""
    import struct

    import numpy as np

    from doubleeye_amd.stream import deir_packet, parse_dedd
    img = np.arange(12, dtype=np.uint8).reshape(3, 4)
    p = deir_packet(img, 1, 77)
    assert p[:4] == b"DEIR" and struct.unpack_from("<HHBBI", p, 4) == (4, 3, 1, 0, 77)
    assert p[14:] == img.tobytes() and len(p) == 14 + 12
    q4 = np.array([[16, -32768, 5, 0]] * 3, np.int16)
    conf = np.full((3, 4), 9, np.uint8)
    pkt = b"DEDD" + struct.pack("<HHII", 4, 3, 77, 1) + img.tobytes() + q4.tobytes() + conf.tobytes()
    (d,) = parse_dedd(pkt)
    assert d["frame"] == 77 and np.array_equal(d["left"], img)
    assert np.array_equal(d["q4"], q4) and np.array_equal(d["conf"], conf)


def test_host_free_rejects_foreign_pointer():
    """de_host_free takes only what de_host_alloc returned; the check is on the
    host, before any HIP call."""
    import numpy as np

    from doubleeye_amd import _lib
    buf = np.zeros(16, np.uint8)
    assert _lib.lib().de_host_free(buf.ctypes.data) == -1      # DE_ERR_ARG
    assert _lib.lib().de_host_free(None) == 0
