"""The large film (BASELINE config 4 as this engine builds it): 200,000
waters + 4,000 surfactants (716,000 sites) as a film 400 x 400 A wide on a
two-layer graphite slab, in the 36^3-box domain (nb2 = 9: the smallest cube of
the reference's 12 A boxes that holds the film), p = 8.

FP64 against the oracle on the whole system (thermal displacement as in the
other full-size checks): classes bit-exact, total force <= 1e-10 per site.
Mixed precision against FP64: RMS relative force error <= 1e-5 at the start,
and again after 100 mixed-precision steps (the "FP64 check every 100 steps"),
every force finite, no error flag raised.
"""
import numpy as np
import pytest

from conftest import _thermal, rel_err

pytestmark = pytest.mark.gpu

NB2, P = 9, 8


def _rms(a, b):
    return float(np.sqrt(((a - b) ** 2).sum() / (b ** 2).sum()))


@pytest.fixture(scope='module')
def c4():
    from surfactant_enhanced_spreading_amd import builder
    return builder.c4_system()


def test_c4_fp64_against_oracle(c4):
    import oracle as orc
    from surfactant_enhanced_spreading_amd import engine
    s = _thermal(c4.copy(), 13)
    print('c4 sites', s.n_sites, flush=True)
    e = engine.Engine(s, nb2=NB2, p=P)
    try:
        e.rebuild()
        e.fmm()
        fe = e.forces()
        ce = e.classes()
    finally:
        e.close()
    o = orc.Oracle(s, nb2=NB2, p=P)
    o.rebuild()
    co = o.classes()
    for k in ('cls', 'anc', 'cbox', 'rkind', 'isub'):
        assert np.array_equal(co[k], ce[k]), k
    o.fmm()
    fo = o.forces()
    err = rel_err(fe, fo)
    print('c4 fp64 total force rel err', err, flush=True)
    assert err <= 1e-10


def test_c4_mixed_tracks_fp64(c4):
    from surfactant_enhanced_spreading_amd import engine
    e32 = engine.Engine(c4, nb2=NB2, p=P, precision=engine.MIXED)
    e64 = engine.Engine(c4, nb2=NB2, p=P)
    try:
        for e in (e32, e64):
            e.rebuild()
            e.fmm()
        f32, f64 = e32.forces(), e64.forces()
        rms0 = _rms(f32, f64)
        e32.bootstrap()
        kb = e32.steps(100, 10, 0)
        assert kb == 100
        r, v, a, eta = e32.state()
        assert np.isfinite(r).all() and np.isfinite(v).all()
        e64.set_state(r, v, a, eta)
        for e in (e32, e64):
            e.rebuild()
            e.fmm()
        f32, f64 = e32.forces(), e64.forces()
        assert np.isfinite(f32).all() and np.isfinite(f64).all()
        rms100 = _rms(f32, f64)
        flags = e32.flags()
        print('c4 mixed rms at start', rms0, 'after 100 steps', rms100, 'flags', flags, flush=True)
        assert rms0 <= 1e-5 and rms100 <= 1e-5
        assert flags['nonfinite'] == 0 and flags['outside'] == 0 and flags['fixed_range'] == 0
    finally:
        e32.close()
        e64.close()
