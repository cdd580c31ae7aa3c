"""The batched ensemble (sesmd_ensemble_*) against its members stepped alone.

A step of the ensemble runs every member's part A, one persistent near-field
launch over all members' row items and every member's site terms and part B,
captured as one graph; with every member at one expansion order the parts
A, B and the site terms of all members are one launch each (batched),
otherwise they run per member.  Bar: bit-for-bit.  Each member's positions,
velocities, accelerations and thermostat variable after fourteen steps (a
one-step graph, a whole ten-step cycle graph with its rebuild, an eager
rebuild step, two more one-step graphs) equal those of an identical context
stepped by sesmd_steps alone, in FP64 and in mixed precision, for the launch
shape the autotuner picks and again after switching to another shape.
Droplets of the ensemble workload's kind (graphite slab three diameters wide,
domain by size), smaller so the test runs in seconds.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _same(a, b):
    ra, va, aa, ea = a.state()
    rb, vb, ab, eb = b.state()
    return (np.array_equal(ra, rb) and np.array_equal(va, vb) and np.array_equal(aa, ab)
            and ea == eb)


@pytest.mark.parametrize('precision,orders', [('fp64', (8,)), ('mixed', (8,)), ('fp64', (8, 4))])
def test_ensemble_bitwise_vs_single(precision, orders):
    """orders (8,): every member at p = 8, so the integrator and site-term
    launches are batched; (8, 4): members alternate between p = 8 and 4, so
    those launches run per member on the branch streams."""
    from surfactant_enhanced_spreading_amd import builder, engine
    prec = {'fp64': engine.FP64, 'mixed': engine.MIXED}[precision]
    systems = builder.ensemble_systems(n=5, n_min=200, n_max=2500, seed=11)
    print('ensemble', [(s.n_water, s.n_surf, nb2) for s, nb2 in systems], 'orders', orders, flush=True)
    solo, team = [], []
    ens = None
    try:
        for k, (s, nb2) in enumerate(systems):
            for lst in (solo, team):
                e = engine.Engine(s, nb2=nb2, p=orders[k % len(orders)], precision=prec)
                e.bootstrap()
                lst.append(e)
        ens = engine.Ensemble(team)
        assert ens.batched() == (len(orders) == 1)
        shape, ms = ens.tune(reps=1)
        print(precision, 'tuned shape', shape, 'ms', ms, flush=True)
        assert ens.shape() == shape
        kb = ens.steps(14, 10, 8)
        assert kb == 22
        for e in solo:
            assert e.steps(14, 10, 8) == 22
        for k, (a, b) in enumerate(zip(team, solo)):
            assert _same(a, b), ('member', k)
        # another launch shape: the same forces, so the same trajectory
        other = (1, 64, 97) if shape != (1, 64, 97) else (1, 128, 1000)
        if prec == engine.MIXED:
            other = (4, 64, 33) if shape != (4, 64, 33) else (2, 128, 500)
        ens.set_shape(*other)
        assert ens.shape() == other
        ens.steps(10, 10, kb)
        for e in solo:
            e.steps(10, 10, kb)
        for k, (a, b) in enumerate(zip(team, solo)):
            assert _same(a, b), ('member after shape change', k)
            assert a.flags() == b.flags()
    finally:
        if ens is not None:
            ens.close()
        for e in solo + team:
            e.close()


def test_ensemble_rejects_bad_input():
    from surfactant_enhanced_spreading_amd import builder, engine
    s, nb2 = builder.ensemble_systems(n=1, n_min=200, n_max=300, seed=3)[0]
    a = engine.Engine(s, nb2=nb2, p=8)
    b = engine.Engine(s, nb2=nb2, p=8, precision=engine.MIXED)
    try:
        with pytest.raises(engine.SesmdError, match='not built'):
            engine.Ensemble([a])
        a.bootstrap()
        b.bootstrap()
        with pytest.raises(engine.SesmdError, match='precision'):
            engine.Ensemble([a, b])
        with pytest.raises(engine.SesmdError, match='repeated'):
            engine.Ensemble([a, a])
        ens = engine.Ensemble([a])
        with pytest.raises(engine.SesmdError, match='one wavefront'):
            ens.set_shape(2, 64, 100)
        with pytest.raises(engine.SesmdError, match='receivers'):
            ens.set_shape(1, 32, 100)
        ens.close()
    finally:
        a.close()
        b.close()
