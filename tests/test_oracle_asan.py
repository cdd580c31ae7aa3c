"""The oracle under AddressSanitizer and UndefinedBehaviorSanitizer: a driver
over every entry point of oracle/sesmd_oracle.h (oracle/asan_driver.c, built
by `make -C oracle asan`) runs the small parity system -- two droplets, one
on the corner shared by eight of the reference's ranks, one in the graphite
layer -- with the reference's graphite sheet, in the reference geometry (p = 4)
and in a 16^3-box domain at p = 8: rebuild, far field, all forces, a masked
near field, init_accel, both integrator halves, 12 steps with a rebuild,
energies, classes and expansions.  Any invalid access or undefined operation
aborts it.  Host code only: no GPU."""
import os
import subprocess

import numpy as np

from conftest import small_system

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_oracle_under_sanitizers(tmp_path):
    from surfactant_enhanced_spreading_amd import formats
    r = subprocess.run(['make', '-s', '-C', os.path.join(ROOT, 'oracle'), 'asan'], capture_output=True, text=True,
                       timeout=600)
    assert r.returncode == 0, r.stderr
    s = small_system()
    g = formats.make_graphite()
    path = str(tmp_path / 'system.bin')
    with open(path, 'wb') as fh:
        np.array([s.n_water, s.n_surf, len(g)], dtype=np.int32).tofile(fh)
        np.ascontiguousarray(s.pos, dtype=np.float64).tofile(fh)
        np.ascontiguousarray(s.vel, dtype=np.float64).tofile(fh)
        np.ascontiguousarray(g, dtype=np.float64).tofile(fh)
    env = dict(os.environ, ASAN_OPTIONS='detect_leaks=0:abort_on_error=1', UBSAN_OPTIONS='halt_on_error=1',
               OMP_NUM_THREADS='1')
    r = subprocess.run([os.path.join(ROOT, 'oracle', '_ref', 'asan_driver'), path], capture_output=True, text=True,
                       timeout=900, env=env)
    print(r.stdout)
    assert r.returncode == 0, r.stderr[-4000:]
    assert 'AddressSanitizer' not in r.stderr and 'runtime error' not in r.stderr, r.stderr[-4000:]
    assert 'asan driver ok' in r.stdout
