"""Captured step graphs against step-by-step stepping.  sesmd_steps replays a
whole rebuild cycle (the rebuild step with its box sort, far field and work
list, then ifar - 1 regular steps) as one graph; stepping one step per call
runs the rebuild step eagerly and each regular step from a one-step graph.
Both issue the same kernels in the same order, so after several cycles --
each replay separated by host reads of the state -- positions, velocities,
accelerations and eta must agree bit for bit, and a masked force evaluation
after a replayed rebuild must see that rebuild's box sort."""
import numpy as np
import pytest

from conftest import medium_system

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize('precision', [0, 1])
def test_cycle_graph_matches_single_steps(precision):
    from surfactant_enhanced_spreading_amd import engine
    s = medium_system(seed=11)
    print('cycle graph vs single steps, precision', precision, s.n_sites, flush=True)
    a = engine.Engine(s, precision=precision)
    b = engine.Engine(s, precision=precision)
    try:
        a.bootstrap()
        b.bootstrap()
        ka = kb = 9  # the next step is a rebuild step: whole cycles from here
        for cycle in range(4):
            ka = a.steps(10, 10, ka)
            for _ in range(10):
                kb = b.steps(1, 10, kb)
            assert ka == kb
            sa, sb = a.state(), b.state()
            for x, y in zip(sa[:3], sb[:3]):
                assert np.array_equal(x, y), cycle
            assert sa[3] == sb[3]
            # a masked evaluation (one reference kernel) after each replayed
            # rebuild: its box-centred list must follow that rebuild
            fa = a.forces(what=engine.NEAR, mask=1 << (cycle + 1))
            fb = b.forces(what=engine.NEAR, mask=1 << (cycle + 1))
            assert np.array_equal(fa, fb), cycle
    finally:
        a.close()
        b.close()
