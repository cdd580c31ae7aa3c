"""Start-up velocities and temperature diagnostics (velocities.py).

The reference's generator is pinned bit-exactly against the compiled
src/random.f90 (live where the reference is built, else against the committed
sequence).  init_temp itself calls MPI and cannot be compiled here; its
replay is pinned by what the routine guarantees: unit vectors drawn in the
reference's storage order from a per-rank restart of the generator, zero
mean per site group on every rank, the target temperature afterwards."""
import json
import os

import numpy as np
import pytest

import oracle as orc
from conftest import small_system, medium_system
from surfactant_enhanced_spreading_amd import forcefield as ff
from surfactant_enhanced_spreading_amd import velocities as V

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLD = json.load(open(os.path.join(ROOT, 'tests', 'golden', 'velocities.json')))
REF_LIB = os.path.join(ROOT, 'oracle', '_ref', 'libmybox_ref.so')


def test_lcg_matches_reference_sequence():
    x = GOLD['seed']
    for h in GOLD['lcg']:
        x = V.reference_lcg(x)
        assert x == float.fromhex(h)


@pytest.mark.skipif(not os.path.exists(REF_LIB), reason='reference Fortran not built here')
def test_lcg_matches_compiled_random_live():
    import ctypes
    import refharness as rh
    f = rh.RefLib().lib.random_
    f.argtypes = [ctypes.POINTER(ctypes.c_double)] * 2
    rng = np.random.default_rng(5)
    for x in np.concatenate([rng.uniform(0, 1, 2000), [V.LCG_SEED, 1e-9, 1 - 1e-9]]):
        a, b = ctypes.c_double(x), ctypes.c_double()
        f(ctypes.byref(a), ctypes.byref(b))
        assert V.reference_lcg(float(x)) == b.value


@pytest.fixture(scope='module')
def filed():
    s = medium_system()
    o = orc.Oracle(s)
    o.rebuild()
    c = o.classes()
    return s, c['cls'], c['anc']


def test_init_temp_replay(filed):
    s, cls, anc = filed
    v = V.init_temp(s, cls, anc)
    # the target temperature with the reference's group masses and df
    assert V.temperature(s, v) == pytest.approx(ff.TR_DEFAULT, rel=1e-13)
    grp = V.site_groups(s.n_water, s.n_surf)
    ranks = V._storage_order(s.n_water, s.n_surf, cls, anc, 12)
    first = np.concatenate([3 * np.arange(s.n_water), 3 * s.n_water + 29 * np.arange(s.n_surf)])
    size = np.concatenate([np.full(s.n_water, 3), np.full(s.n_surf, 29)])
    nonempty = 0
    scales = []
    for mols in ranks:
        if len(mols) == 0:
            continue
        nonempty += 1
        sites = np.concatenate([np.arange(first[m], first[m] + size[m]) for m in mols])
        # zero mean per group on every rank
        for g in range(5):
            sel = sites[grp[sites] == g]
            if len(sel):
                assert np.abs(v[sel].sum(axis=0)).max() <= 1e-12 * np.abs(v[sel]).sum()
        # the rank's first site: the first three draws after the seed, as a unit
        # vector, minus the rank's group mean, times the global factor
        x, d = V.LCG_SEED, []
        for _ in range(3):
            x = V.reference_lcg(x)
            d.append(2.0 * (x - 0.5))
        u = np.array(d) / np.linalg.norm(d)
        i0 = sites[0]
        sel = sites[grp[sites] == grp[i0]]
        # undo: v = scale (u_i - mean); mean from the rank's raw vectors
        raw = np.zeros((len(sites), 3))
        x = V.LCG_SEED
        for k in range(len(sites)):
            dd = []
            for _ in range(3):
                x = V.reference_lcg(x)
                dd.append(2.0 * (x - 0.5))
            raw[k] = np.array(dd) / np.linalg.norm(dd)
        assert np.allclose(raw[0], u, rtol=0, atol=1e-15)
        pos = {int(sv): k for k, sv in enumerate(sites)}
        w = u - raw[[pos[int(t)] for t in sel]].mean(axis=0)
        if np.linalg.norm(w) < 1e-12:  # a group of one site: zero after the mean
            assert np.abs(v[i0]).max() <= 1e-15
            continue
        k = int(np.argmax(np.abs(w)))
        scales.append(v[i0][k] / w[k])
        assert np.allclose(v[i0], scales[-1] * w, rtol=1e-12, atol=1e-15)
    assert nonempty >= 4 and len(scales) >= 2
    # one global factor sets the temperature for every rank
    assert np.ptp(scales) <= 1e-12 * abs(scales[0])


def test_storage_order_is_box_class_molecule(filed):
    s, cls, anc = filed
    for mols in V._storage_order(s.n_water, s.n_surf, cls, anc, 12):
        a = anc[mols]
        ib4 = (a % 12 % 4) + 4 * ((a // 12) % 12 % 4) + 16 * (a // 144 % 4)
        key = list(zip(ib4, cls[mols], mols))
        assert key == sorted(key)


def test_projection_and_maxwell_boltzmann():
    s = small_system()
    v = V.maxwell_boltzmann(s, seed=3)
    assert V.temperature(s, v) == pytest.approx(ff.TR_DEFAULT, rel=1e-12)
    ia, ib, _ = V._bonds(s)
    for a, b in zip(ia, ib):
        rab = s.pos[a] - s.pos[b]
        vab = v[a] - v[b]
        rv = np.abs((rab * vab).sum(axis=1))
        assert rv.max() <= 1e-12 * np.abs(vab).max() * np.abs(rab).max()
    # idempotent, and it only removes kinetic energy
    w = V.project_constraints(s, v)
    assert np.abs(w - v).max() <= 1e-14 * np.abs(v).max()
    rng = np.random.default_rng(9)
    raw = rng.normal(size=v.shape) * 1e-3
    pr = V.project_constraints(s, raw)
    m = s.masses()
    assert (m[:, None] * pr ** 2).sum() < (m[:, None] * raw ** 2).sum()


def test_temperature_and_counts_match_oracle(filed):
    s, cls, anc = filed
    o = orc.Oracle(s)
    o.rebuild()
    e = o.energies()
    assert V.temperature(s) == pytest.approx(e['temperature'], rel=1e-12)
    c = V.molecule_counts(cls)
    assert c['w1'] + c['w2'] == s.n_water and c['surfactant'] == s.n_surf
    assert c['sites'] == s.n_sites
