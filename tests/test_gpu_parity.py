"""HIP engine (libsesmd.so) against the C restatement of the reference, FP64.

Bar: integer rebuild outputs bit-exact; per-site relative force error
(floored at 1e-3 of the RMS force) <= 1e-10; one full step <= 1e-10.
"""
import numpy as np
import pytest

from conftest import rel_err

pytestmark = pytest.mark.gpu

TOL = 1e-10


@pytest.fixture(scope='module', params=['small', 'medium'])
def pair(request, sys_small, sys_medium):
    import oracle as orc
    from surfactant_enhanced_spreading_amd import engine
    s = sys_small if request.param == 'small' else sys_medium
    print('system', request.param, s.n_sites, flush=True)
    o = orc.Oracle(s)
    o.rebuild()
    o.fmm()
    e = engine.Engine(s)
    e.rebuild()
    e.fmm()
    yield o, e
    e.close()


def test_rebuild_bit_exact(pair):
    o, e = pair
    a, b = o.classes(), e.classes()
    for k in ('cls', 'anc', 'cbox', 'rkind', 'isub'):
        assert np.array_equal(a[k], b[k]), k


def test_expansions(pair):
    o, e = pair
    A, B = o.expansions(), e.expansions()
    for k in ('O4', 'O3', 'O2', 'L2', 'L3', 'L4'):
        scale = np.abs(A[k]).max()
        assert scale > 0, k
        err = np.abs(A[k] - B[k]).max() / scale
        assert err < 1e-12, (k, err)


@pytest.mark.parametrize('what', ['near', 'far', 'internal', 'subtract', 'valence',
                                  'dihedral', 'dihedral_be', 'wall', 'graphite', 'all'])
def test_force_component(pair, what):
    import oracle as orc
    from surfactant_enhanced_spreading_amd import engine
    o, e = pair
    bit = {'near': 1, 'far': 2, 'internal': 4, 'subtract': 8, 'valence': 16, 'dihedral': 32,
           'dihedral_be': 64, 'wall': 128, 'graphite': 256, 'all': 511}[what]
    fo = o.forces(bit)
    fe = e.forces(bit)
    if what == 'graphite':
        assert np.abs(fo).max() > 0
    assert rel_err(fe, fo) <= TOL, what


def test_reference_kernel_masks(pair):
    """Each of the 39 eval[wS]* kernels alone (the drop-in granularity)."""
    import oracle as orc
    o, e = pair
    for name, pairs in orc.KERNEL_PAIRS.items():
        m = orc.near_mask(pairs)
        fo = o.forces(orc.NEAR, m)
        fe = e.forces(orc.NEAR, m)
        if np.abs(fo).max() == 0:
            assert np.abs(fe).max() == 0, name
        else:
            assert rel_err(fe, fo) <= TOL, name


def test_one_step(pair, sys_small, sys_medium):
    """bootstrap + init_accel + part A + rebuild + forces + part B."""
    import oracle as orc
    from surfactant_enhanced_spreading_amd import engine
    o, e = pair
    f0 = o.forces()
    o.init_accel(f0)
    e.forces(fetch=False)
    e.init_accel()
    ro, vo, ao, eo = o.state()
    re_, ve, ae, ee = e.state()
    assert rel_err(ae, ao) <= TOL
    so = o.move_a()
    se = e.move_a()
    assert np.abs(se - so).max() <= 1e-12 * np.abs(so).max()
    ro, vo, ao, eo = o.state()
    re_, ve, ae, ee = e.state()
    assert np.abs(re_ - ro).max() <= 1e-12 * np.abs(ro).max()
    assert rel_err(ve, vo) <= TOL
    assert abs(ee - eo) <= 1e-12 * abs(eo)
    o.rebuild(); o.fmm()
    e.rebuild(); e.fmm()
    f1 = o.forces()
    fe1 = e.forces()
    assert rel_err(fe1, f1) <= TOL
    o.move_b(f1)
    e.move_b()
    ro, vo, ao, eo = o.state()
    re_, ve, ae, ee = e.state()
    assert rel_err(ve, vo) <= TOL
    assert rel_err(ae, ao) <= TOL


@pytest.mark.parametrize('which,ifar', [('small', 10), ('medium', 10), ('small', 1)])
def test_steps_pipeline(which, ifar, sys_small, sys_medium):
    """sesmd_steps -- steps replayed from captured graphs (one regular step,
    and a whole rebuild cycle: the rebuild + far-field step and its ifar - 1
    regular steps), partial cycles with an eager rebuild step -- against the
    oracle's own step loop.  ifar = 10: steps 9..22 (step 9 from the one-step
    graph, 10..19 as one cycle graph, 20 an eager rebuild, 21..22 from the
    one-step graph); ifar = 1: steps 1..5, each a one-step cycle graph."""
    import oracle as orc
    from surfactant_enhanced_spreading_amd import engine
    s = sys_small if which == 'small' else sys_medium
    o = orc.Oracle(s)
    o.rebuild()
    o.fmm()
    o.init_accel(o.forces())
    e = engine.Engine(s)
    try:
        e.bootstrap()
        n, k0 = (14, 8) if ifar == 10 else (5, 0)
        ko = o.steps(n, ifar, k0)
        ke = e.steps(n, ifar, k0)
        assert ko == ke == n + k0
        ro, vo, ao, eo = o.state()
        re_, ve, ae, ee = e.state()
        print(which, 'r', np.abs(re_ - ro).max(), 'v', rel_err(ve, vo), 'a', rel_err(ae, ao),
              'eta', ee, eo, flush=True)
        assert np.abs(re_ - ro).max() <= 1e-12 * np.abs(ro).max()
        assert rel_err(ve, vo) <= TOL
        assert rel_err(ae, ao) <= TOL
        assert abs(ee - eo) <= 1e-10 * abs(eo)
        assert e.shake_failures() == o.shake_failures()
    finally:
        e.close()


@pytest.mark.parametrize('nb2,p', [(3, 8), (4, 8), (4, 4), (5, 4)])
def test_other_orders_and_domains(nb2, p, sys_small):
    """The engine beyond the reference's own p = 4, 27-rank geometry: expansion
    order 8, a 16^3-box domain (nb2 = 4) and a 20^3-box one (nb2 = 5: more
    rows than the row list keeps in LDS, and odd half-widths at L3), against
    the oracle run with the same parameters -- classes bit-exact, expansions
    1e-12, total force 1e-10."""
    import oracle as orc
    from surfactant_enhanced_spreading_amd import engine
    o = orc.Oracle(sys_small, nb2=nb2, p=p)
    o.rebuild()
    o.fmm()
    e = engine.Engine(sys_small, nb2=nb2, p=p)
    try:
        e.rebuild()
        e.fmm()
        a, b = o.classes(), e.classes()
        for k in ('cls', 'anc', 'cbox', 'rkind', 'isub'):
            assert np.array_equal(a[k], b[k]), k
        A, B = o.expansions(), e.expansions()
        for k in ('O4', 'O3', 'O2', 'L2', 'L3', 'L4'):
            scale = np.abs(A[k]).max()
            assert scale > 0, k
            assert np.abs(A[k] - B[k]).max() <= 1e-12 * scale, k
        assert rel_err(e.forces(), o.forces()) <= TOL
    finally:
        e.close()
