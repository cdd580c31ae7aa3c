"""The C ABI library loads and exports every function include/sesmd.h declares."""
import ctypes
import os
import re
import subprocess
import sys

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


def test_fortran_library_exports_reference_names():
    """libsesmd_fortran.so exports every routine include/sesmd_fortran.h
    declares under the reference's Fortran external names, and loads without
    the COMMON blocks (weak references, resolved by the Fortran program).
    Checked in a child process: a library loaded without COMMON keeps its
    unresolved references for the life of the process."""
    txt = open(os.path.join(ROOT, 'include', 'sesmd_fortran.h')).read()
    txt = re.sub(r'/\*.*?\*/', '', txt, flags=re.S)
    names = sorted(set(re.findall(r'\b([a-z0-9_]+_|sesmd_fortran_[a-z_]+)\s*\(', txt)))
    fortran = [n for n in names if n.endswith('_')]
    assert len(fortran) >= 39 + 8 + 9 + 7 + 10
    path = os.path.join(ROOT, 'surfactant_enhanced_spreading_amd', 'libsesmd_fortran.so')
    code = """
import ctypes, sys
L = ctypes.CDLL(sys.argv[1])
for n in sys.argv[2:]:
    assert hasattr(L, n), n
L.sesmd_fortran_last_error.restype = ctypes.c_char_p
before = L.sesmd_fortran_errors()
L.evalw1w1_(ctypes.byref(ctypes.c_int(0)))
assert L.sesmd_fortran_errors() == before + 1
assert b'COMMON' in L.sesmd_fortran_last_error()
print('ok')
"""
    r = subprocess.run([sys.executable, '-c', code, path] + names, capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 0 and r.stdout.strip() == 'ok', r.stderr
