"""The C restatement against the reference's compiled Fortran, run live (needs
oracle/_ref/libmybox_ref.so, i.e. a checkout with the reference tree; skipped
elsewhere -- the committed fixtures cover the same ground there)."""
import os

import numpy as np
import pytest

from conftest import rel_err, _thermal
from surfactant_enhanced_spreading_amd import builder

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
pytestmark = pytest.mark.skipif(
    not os.path.exists(os.path.join(ROOT, 'oracle', '_ref', 'libmybox_ref.so')),
    reason='reference Fortran not built here (make -C oracle ref)')


@pytest.fixture(scope='module')
def pair():
    import oracle as orc
    import refharness as rh
    # a different layout from the fixtures: a droplet astride an L2 face, one
    # near the -z wall in the graphite layer, one against the +y wall
    d1 = builder.droplet(220, 4, center_A=(-24.0, 5.0, 10.0), seed=21)
    d2 = builder.droplet(45, 1, center_A=(30.0, -20.0, -63.0), seed=22)
    d3 = builder.droplet(40, 1, center_A=(0.0, 62.0, 0.0), seed=23)
    s = _thermal(builder.assemble([d1, d2, d3], graphite=builder.graphite_sheet(), seed=21), 21)
    ref = rh.RefLib()
    em = rh.Emulator(ref, s)
    o = orc.Oracle(s)
    o.rebuild()
    o.fmm()
    return em, o, s


def test_integer_maps(pair):
    em, o, s = pair
    c = o.classes()
    assert np.array_equal(c['cls'], em.cls)
    assert np.array_equal(c['anc'], em.anc[:, 0] + 12 * (em.anc[:, 1] + 12 * em.anc[:, 2]))
    assert np.array_equal(c['isub'], em.isub_masks())


def test_total_force(pair):
    em, o, s = pair
    assert rel_err(o.forces(), em.forces()) <= 1e-10


def test_kernels(pair):
    import oracle as orc
    em, o, s = pair
    for k, pr in orc.KERNEL_PAIRS.items():
        fr = em.forces(far=False, intra=False, wall=False, graphite=False, kernels=[k])
        fo = o.forces(orc.NEAR, orc.near_mask(pr))
        if np.abs(fr).max() == 0:
            assert np.abs(fo).max() == 0, k
        else:
            assert rel_err(fo, fr) <= 1e-10, k


def test_two_steps(pair):
    em, o, s = pair
    o.init_accel(o.forces())
    em.init_accel(em.forces())
    for _ in range(2):
        em.step_a()
        o.move_a()
        em.reclassify()
        o.rebuild()
        o.fmm()
        f = o.forces()
        fr = em.forces()
        assert rel_err(f, fr) <= 1e-10
        em.step_b(fr)
        o.move_b(f)
    r, v, a, eta = o.state()
    assert np.abs(r - em.pos).max() <= 1e-11 * np.abs(em.pos).max()
    assert rel_err(v, em.vel) <= 1e-10
    assert abs(eta - em.eta) <= 1e-11 * abs(em.eta)
