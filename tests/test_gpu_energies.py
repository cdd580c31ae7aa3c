"""sesmd_energies on the GPU against the oracle's energies (which are pinned
to the forces by finite differences, tests/test_energies.py): every component
within 1e-10 of the scale of the energies, at the bootstrap state and after
a run of steps with rebuilds (thermostat integral included)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

COMPS = ('near_coul', 'near_lj', 'far', 'subtract', 'internal_coul', 'internal_lj', 'angle',
         'dihedral', 'dihedral_be', 'wall', 'graphite', 'potential', 'kinetic', 'thermostat',
         'total', 'virial', 'temperature')


def _compare(eg, eo, tol):
    scale = sum(abs(eo[k]) for k in COMPS[:11]) + abs(eo['kinetic']) + abs(eo['thermostat'])
    for k in COMPS:
        ref = abs(eo[k]) if k == 'temperature' else scale
        assert abs(eg[k] - eo[k]) <= tol * ref, (k, eg[k], eo[k])


@pytest.mark.parametrize('which', ['small', 'medium'])
def test_energies_match_oracle(which, sys_small, sys_medium):
    import oracle as orc
    from surfactant_enhanced_spreading_amd import engine
    s = sys_small if which == 'small' else sys_medium
    print('energies', which, s.n_sites, flush=True)
    o = orc.Oracle(s)
    o.rebuild()
    o.fmm()
    e = engine.Engine(s)
    try:
        e.rebuild()
        e.fmm()
        _compare(e.energies(), o.energies(), 1e-10)
        o.init_accel(o.forces())
        e.bootstrap()
        ko = o.steps(25, 10, 3)
        ke = e.steps(25, 10, 3)
        assert ko == ke
        eo, eg = o.energies(), e.energies()
        print(which, {k: (eg[k], eo[k]) for k in ('total', 'thermostat', 'kinetic')}, flush=True)
        assert eo['thermostat'] != 0.0
        _compare(eg, eo, 1e-9)
    finally:
        e.close()


def test_mixed_context_energies_are_fp64(sys_small):
    """A mixed-precision context evaluates energies in FP64: at the same state
    they equal the FP64 context's, except the far-field term, which reads the
    local expansions that the mixed context's split-BF16x3 M2L produced (FP32
    accumulation: relative error of order 1e-7, bound 1e-5 here)."""
    from surfactant_enhanced_spreading_amd import engine
    e64 = engine.Engine(sys_small)
    em = engine.Engine(sys_small, precision=engine.MIXED)
    try:
        for e in (e64, em):
            e.rebuild()
            e.fmm()
        a, b = e64.energies(), em.energies()
        d = a['far'] - b['far']
        assert a['far'] != 0.0 and abs(d) <= 1e-5 * abs(a['far']), (a['far'], b['far'])
        for k in COMPS:
            if k in ('potential', 'total'):  # sums that hold the far term
                assert abs((a[k] - b[k]) - d) <= 1e-12 * abs(a[k]), k
            elif k != 'far':
                assert a[k] == b[k], k
    finally:
        e64.close()
        em.close()
