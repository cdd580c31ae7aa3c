"""Build libsesmd.so (the HIP engine) and libsesmd_fortran.so (the routines
under the reference's Fortran names, over libsesmd) in-tree for gfx950.

    python -m surfactant_enhanced_spreading_amd.csrc.build
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.dirname(HERE)
ROOT = os.path.dirname(PKG)
OUT = os.path.join(PKG, 'libsesmd.so')
OUT_FORTRAN = os.path.join(PKG, 'libsesmd_fortran.so')
ARCH = os.environ.get('SESMD_ARCH', 'gfx950')

# -packed-fp32-ops off: no v_pk_*_f32 generated from scalar/float2 code.
FLAGS = ['-O3', '-std=c++17', '-fPIC', '-shared', '--offload-arch=' + ARCH,
         '-Xclang', '-target-feature', '-Xclang', '-packed-fp32-ops',
         '-Wno-unused-result', '-Wno-unused-value', '-I' + os.path.join(ROOT, 'include')]


def build(verbose=False):
    from surfactant_enhanced_spreading_amd.csrc import gen_ff_header
    gen_ff_header.write()
    hipcc = os.environ.get('HIPCC', '/opt/rocm/bin/hipcc')
    src = os.path.join(HERE, 'sesmd.hip')
    tmp = OUT + '.tmp'
    cmd = [hipcc] + FLAGS + ['-o', tmp, src]
    if verbose:
        print(' '.join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    err = '\n'.join(l for l in r.stderr.splitlines()
                    if 'packed-fp32-ops' not in l)
    if r.returncode != 0:
        raise RuntimeError('hipcc failed:\n' + err)
    if verbose and err.strip():
        print(err)
    os.replace(tmp, OUT)
    build_fortran(verbose)
    return OUT


def build_fortran(verbose=False):
    """The Fortran-named library: host C++ only, linked against libsesmd.so."""
    from surfactant_enhanced_spreading_amd.csrc import gen_ff_header
    gen_ff_header.write()
    tmp = OUT_FORTRAN + '.tmp'
    cmd = [os.environ.get('CXX', 'g++'), '-O2', '-std=c++17', '-fPIC', '-shared', '-Wall',
           '-I' + os.path.join(ROOT, 'include'), '-I' + HERE, os.path.join(HERE, 'dropin.cpp'),
           '-L' + PKG, '-lsesmd', '-Wl,-rpath,$ORIGIN', '-o', tmp]
    if verbose:
        print(' '.join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError('g++ failed:\n' + r.stderr)
    os.replace(tmp, OUT_FORTRAN)
    return OUT_FORTRAN


if __name__ == '__main__':
    print(build(verbose=True))
