"""The C ABI library loads and exports every function include/sesmd.h declares."""
import ctypes
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared(header):
    txt = open(os.path.join(ROOT, 'include', header)).read()
    txt = re.sub(r'/\*.*?\*/', '', txt, flags=re.S)
    return sorted(set(re.findall(r'\b(sesmd_[a-z0-9_]+)\s*\(', txt)))


def test_library_exports_header_symbols():
    from surfactant_enhanced_spreading_amd import engine
    L = ctypes.CDLL(engine.LIB_PATH)
    names = declared('sesmd.h')
    assert len(names) >= 20
    for n in names:
        assert hasattr(L, n), n
