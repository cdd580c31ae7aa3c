"""P2M pinned to the reference's own computeOnm4 (src/computeOnm4.f90:22-162).

computeOnm4 also calls MPI (three diagnostic reductions, lines 203-207), so
the compiled subset oracle/refharness.py normally loads leaves it out and the
harness restates it (Emulator.p2m, with the reference's own Ylm).  Here the
routine itself runs: oracle/_ref/libmybox_mpi.so (every reference routine,
built by `make -C oracle mpi` against the image's MPICH) is loaded in a
subprocess that calls MPI_Init as a singleton, fills each of the 27 ranks'
storage and lists as the harness does, calls computeonm4_ and reads cO4 --
for every rank whose own boxes hold a water oxygen: with one process the
routine's diagnostic O:H ratio (line 211) divides by that rank's own oxygen
count.  Bar: every own box of those ranks within 1e-13 of the largest
coefficient of the
restatement (the same sums; only the powers rho^n are formed differently).
The restatement is in turn what the oracle's and the engine's P2M are pinned
against, to 1e-12 (test_oracle_reference.py, test_gpu_parity.py).
"""
import json
import os
import subprocess
import sys

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
LIB = os.path.join(ROOT, 'oracle', '_ref', 'libmybox_mpi.so')

SCRIPT = r'''
import ctypes, json, sys
import numpy as np
sys.path[:0] = [%r, %r, %r]
import refharness as rh
from conftest import medium_system
ref = rh.RefLib(path=%r)
mpi = ctypes.CDLL('libmpi.so.12', mode=ctypes.RTLD_GLOBAL)
assert mpi.MPI_Init(None, None) == 0
em = rh.Emulator(ref, medium_system())
O4 = em.p2m()
worst, boxes = 0.0, 0
scale = np.abs(O4).max()
ranks = 0
for r, rk in enumerate(em.ranks()):
    em.load_rank(rk)
    em.build_lists(r)
    # oxygens the routine counts in this rank's own boxes (W1 molecules, W2
    # central sites of type 1; src/computeOnm4.f90:36-73): with one process
    # its diagnostic ratio ihgntotal / iogntotal divides by this count
    ox = 0
    for i in range(64):
        ox += int(ref['ictw1'][em._ib(i, 6, 1) - 1])
        n2 = int(ref['ictW2Central'][i])
        ox += int((ref['W2Central'][:n2, 3, i] == 1).sum())
    if ox == 0:
        continue
    ranks += 1
    ref.call('computeonm4', r)
    c = ref['cO4']
    for i in range(64):
        g = em._glob(rk, i)
        got = c[:, em._ib(i, 8, 2) - 1]
        worst = max(worst, float(np.abs(got - O4[g[2], g[1], g[0]]).max()))
        boxes += 1
mpi.MPI_Finalize()
print(json.dumps(dict(worst=worst, scale=float(scale), boxes=boxes, ranks=ranks)))
'''


@pytest.mark.skipif(not os.path.exists(LIB), reason='reference MPI build absent (make -C oracle mpi)')
def test_compiled_computeOnm4_matches_restatement(tmp_path):
    code = SCRIPT % (os.path.join(ROOT, 'oracle'), os.path.join(ROOT, 'tests'), ROOT, LIB)
    r = subprocess.run([sys.executable, '-c', code], cwd=str(tmp_path), capture_output=True,
                       text=True, timeout=900)
    assert r.returncode == 0, r.stderr[-2000:]
    res = json.loads(r.stdout.strip().splitlines()[-1])
    print(res)
    assert res['ranks'] >= 8 and res['boxes'] == 64 * res['ranks']
    assert res['scale'] > 0.0
    assert res['worst'] <= 1e-13 * res['scale'], res
