"""The HIP engine against the reference's own compiled routines on its
workload of record: 9,997 waters + 475 surfactants (src/mybox.cmns:59-60) in
the reference geometry (12^3 L4 boxes, p = 4), read as the data_cont2.dat the
reference program reads (tests/golden/r2_system.npz, written by
tests/golden/make_goldens.py r2 through oracle/refharness.py).

Bar, at every 7th site: bootstrap total force within 1e-10 (per-site relative
error floored at 1e-3 of the RMS force); after one full step (part A, the
rebuild and far field, forces, part B) positions within 1e-12 of their
largest coordinate, the thermostat variable within 1e-12, velocities and
accelerations within 1e-9 (the step tests' bar: one force evaluation at the
1e-10 bar fed through SHAKE/RATTLE).
"""
import numpy as np
import pytest

from conftest import rel_err
from test_oracle_golden import load, r2_system_from_golden

pytestmark = pytest.mark.gpu


def test_r2_against_compiled_reference():
    from surfactant_enhanced_spreading_amd import engine
    z = load('r2_system.npz')
    s = r2_system_from_golden(z)
    k = z['sites']
    print('R2 system', s.n_sites, 'sites', flush=True)
    e = engine.Engine(s, nb2=3, p=4)
    try:
        e.rebuild()
        e.fmm()
        f = e.forces()
        err = rel_err(f[k], z['F'])
        print('bootstrap force rel err', err, flush=True)
        assert err <= 1e-10
        e.init_accel()
        assert e.steps(1, 10, 9) == 10
        r, v, a, eta = e.state()
        print('step: r', np.abs(r[k] - z['A_pos']).max(), 'v', rel_err(v[k], z['B_vel']),
              'a', rel_err(a[k], z['B_acc']), 'eta', eta, float(z['A_eta']), flush=True)
        assert np.abs(r[k] - z['A_pos']).max() <= 1e-12 * np.abs(z['A_pos']).max()
        assert abs(eta - float(z['A_eta'])) <= 1e-12 * abs(float(z['A_eta']))
        assert rel_err(v[k], z['B_vel']) <= 1e-9
        assert rel_err(a[k], z['B_acc']) <= 1e-9
    finally:
        e.close()
