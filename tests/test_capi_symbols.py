"""The C-ABI library loads on a CPU-only host and exports every declared entry point."""
import ctypes
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    names = set()
    inc = os.path.join(ROOT, "include")
    for f in os.listdir(inc):
        if f.endswith(".h"):
            txt = open(os.path.join(inc, f)).read()
            names |= set(re.findall(r"^\s*(?:int|double|void)\s+(nv2_\w+)\s*\(", txt, re.M))
    return names


def test_library_exports_header_symbols():
    from nevpt2_amd import _lib
    lib = ctypes.CDLL(_lib.LIBPATH)
    names = declared()
    assert len(names) >= 6
    missing = [n for n in sorted(names) if not hasattr(lib, n)]
    assert not missing, missing


def test_error_channel_without_gpu():
    """Invalid arguments are rejected before any device work, with a readable message."""
    from nevpt2_amd import _lib
    L = _lib.lib()
    rc = L.nv2_bro(0, 1, 6, None, None, None, None)
    assert rc == -2
    assert "nact" in _lib.last_error()


def test_noalloc_rejects_sizes_beyond_its_scratch():
    """The allocation-free entry point refuses nact above its static scratch before any device work."""
    from nevpt2_amd import _lib
    L = _lib.lib()
    rc = L.nv2_koop_matrices_noalloc(1, 17, 2, 20, *([None] * 19), None)
    assert rc == -2
    assert "nact=17" in _lib.last_error() and "scratch" in _lib.last_error()
