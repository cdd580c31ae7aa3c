"""Energies (not in the reference): every component of the oracle's energy is
pinned against its force term by central finite differences, so a dropped
factor, a wrong sign or a missing pair shows up as a mismatch with forces that
are themselves pinned to the reference's routines.  The thermostat integral
and the kinetic terms are checked against their closed forms."""
import zlib

import numpy as np
import pytest

import oracle as orc
from conftest import small_system, medium_system
from surfactant_enhanced_spreading_amd import forcefield as ff

# force bit -> (energy components, factor): -factor * dE/dx must equal the force.
# The far field is 1/2 sum q Phi while the L2P force is -q grad Phi with the
# expansion frozen, hence factor 2.
TERMS = {
    'near': (orc.NEAR, ('near_coul', 'near_lj'), 1.0),
    'far': (orc.FAR, ('far',), 2.0),
    'internal': (orc.INTERNAL, ('internal_coul', 'internal_lj'), 1.0),
    'subtract': (orc.SUBTRACT, ('subtract',), 1.0),
    'valence': (orc.VALENCE, ('angle',), 1.0),
    'dihedral': (orc.DIHEDRAL, ('dihedral',), 1.0),
    'dihedral_be': (orc.DIHEDRAL_BE, ('dihedral_be',), 1.0),
    'wall': (orc.WALL, ('wall',), 1.0),
    'graphite': (orc.GRAPHITE, ('graphite',), 1.0),
}


@pytest.fixture(scope='module')
def medium_oracle():
    s = medium_system()
    o = orc.Oracle(s)
    o.rebuild()
    o.fmm()
    return o, s


def _pick(o, s, what, rng, k=8):
    """Sites where the term is non-zero (the largest few and a random few)."""
    F = o.forces(what)
    mag = np.abs(F).max(axis=1)
    nz = np.nonzero(mag > 0)[0]
    assert len(nz) > 0
    top = nz[np.argsort(mag[nz])[-k // 2:]]
    rnd = rng.choice(nz, min(k // 2, len(nz)), replace=False)
    return F, np.unique(np.concatenate([top, rnd]))


@pytest.mark.parametrize('term', sorted(TERMS))
def test_energy_gradient_is_force(term, medium_oracle):
    o, s = medium_oracle
    what, keys, fac = TERMS[term]
    rng = np.random.default_rng(zlib.crc32(term.encode()) % 1000)  # the same sites every run
    F, sites = _pick(o, s, what, rng)
    r0, v0, a0, eta = o.state()
    h = 1e-6
    try:
        for i in sites:
            scale = max(np.abs(F[i]).max(), 1e-8)
            for a in range(3):
                r = r0.copy()
                r[i, a] += h
                o.set_state(r, v0, a0, eta)
                ep = sum(o.energies()[k] for k in keys)
                r[i, a] -= 2 * h
                o.set_state(r, v0, a0, eta)
                em = sum(o.energies()[k] for k in keys)
                fd = -fac * (ep - em) / (2 * h)
                # central differences of energies of size ~1e3: ~1e-7 absolute noise
                assert abs(fd - F[i, a]) <= 2e-5 * scale + 2e-6, (term, i, a, fd, F[i, a])
    finally:
        o.set_state(r0, v0, a0, eta)


def test_totals_and_closed_forms():
    s = small_system()
    o = orc.Oracle(s)
    o.rebuild()
    o.fmm()
    e = o.energies()
    comps = ('near_coul', 'near_lj', 'far', 'subtract', 'internal_coul', 'internal_lj', 'angle',
             'dihedral', 'dihedral_be', 'wall', 'graphite')
    assert e['potential'] == pytest.approx(sum(e[k] for k in comps), rel=1e-14)
    m = s.masses()
    kin = 0.5 * (m[:, None] * s.vel ** 2).sum() / ff.DT ** 2
    assert e['kinetic'] == pytest.approx(kin, rel=1e-12)
    df = 3 * s.n_sites - (3 * s.n_water + 28 * s.n_surf)
    assert e['temperature'] == pytest.approx(2 * kin / df, rel=1e-12)
    assert e['thermostat'] == 0.0  # eta = 0, xi = 0
    # xi is the trapezoid integral of eta over the steps' eta updates
    o.init_accel(o.forces())
    etas = [o.state()[3]]
    for _ in range(3):
        o.move_a()
        etas.append(o.state()[3])
        o.move_b(o.forces())
    xi = sum(0.5 * ff.DT * (etas[k] + etas[k + 1]) for k in range(3))
    assert o.xi() == pytest.approx(xi, rel=1e-14)
    e = o.energies()
    pnh1 = ff.DT / df / ff.TR_DEFAULT / ff.RLXTIME ** 2
    q = ff.DT / (2 * pnh1)  # = df tr tau^2 / 2
    assert e['thermostat'] == pytest.approx(0.5 * q * etas[-1] ** 2 + df * ff.TR_DEFAULT * xi, rel=1e-12)
    # the pair virial of Coulomb pairs alone is their energy (r . F = q q / r)
    assert e['virial'] != 0.0


def test_extended_energy_is_conserved():
    """After the first steps (which project the initial velocities onto the
    constraints and so remove kinetic energy), the extended Nose-Hoover energy
    stays within a few per cent of the kinetic energy over 80 steps, list and
    far-field rebuilds included, while the thermostat exchanges far more."""
    s = small_system()
    o = orc.Oracle(s)
    o.rebuild()
    o.fmm()
    o.init_accel(o.forces())
    kb = o.steps(20, 10, 0)
    e0 = o.energies()
    hs, th = [], []
    for _ in range(8):
        kb = o.steps(10, 10, kb)
        e = o.energies()
        hs.append(e['total'])
        th.append(e['thermostat'])
    dev = np.abs(np.array(hs) - e0['total']).max()
    assert dev < 0.03 * e0['kinetic'], (dev, e0['kinetic'])
    assert np.abs(np.array(th) - e0['thermostat']).max() > 3 * dev
