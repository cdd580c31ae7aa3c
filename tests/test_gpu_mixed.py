"""Mixed precision (SESMD_MIXED: FP32 pair arithmetic with fixed-point
accumulation at 2^-32, M2L as split-BF16x3 MFMA products accumulated in FP32, and
everything else FP64) against the FP64 engine, which is itself checked
against the oracle to 1e-10.

Bar (the accuracy target of the mixed mode): RMS relative force error
||F_mixed - F_fp64||_2 / ||F_fp64||_2 <= 1e-5 over all sites, on the test
systems and on the 65,800-site droplet the benchmark runs; no site worse than
1e-3 of its own force scale (per-site error floored at 1e-3 of the RMS force).
Integer work (classes, boxes, sort) is shared and must be identical.  The
expansions: multipoles and the top level identical (FP64 in both modes); the
M2L levels within 1e-6 of their largest coefficient (three bf16 parts carry
the 24 bits of an FP32 value; six of the nine cross products are kept).
"""
import numpy as np
import pytest

from conftest import rel_err

pytestmark = pytest.mark.gpu

RMS_TOL = 1e-5


def _rms(a, b):
    return float(np.sqrt(((a - b) ** 2).sum() / (b ** 2).sum()))


def _pair(s):
    from surfactant_enhanced_spreading_amd import engine
    e64 = engine.Engine(s)
    e32 = engine.Engine(s, precision=engine.MIXED)
    for e in (e64, e32):
        e.rebuild()
        e.fmm()
    return e64, e32


@pytest.mark.parametrize('which', ['small', 'medium', 'bench'])
def test_mixed_force_error(which, sys_small, sys_medium):
    from surfactant_enhanced_spreading_amd import builder
    s = {'small': sys_small, 'medium': sys_medium}.get(which)
    if s is None:
        s = builder.bench_system()
    print('system', which, s.n_sites, flush=True)
    e64, e32 = _pair(s)
    try:
        c64, c32 = e64.classes(), e32.classes()
        for k in ('cls', 'anc', 'cbox', 'rkind', 'isub'):
            assert np.array_equal(c64[k], c32[k]), k
        f64n, f32n = e64.forces(1), e32.forces(1)
        f64, f32 = e64.forces(511), e32.forces(511)
        rms_near = _rms(f32n, f64n)
        rms = _rms(f32, f64)
        worst = rel_err(f32, f64)
        print(which, 'rms near', rms_near, 'rms total', rms, 'worst site', worst, flush=True)
        assert rms_near > 0.0          # the mixed path is really taken
        assert rms <= RMS_TOL, rms
        assert worst <= 1e-3, worst
        # the masked (per reference kernel) path stays FP64
        assert np.array_equal(e32.forces(1, mask=(1 << 54) - 1), e64.forces(1, mask=(1 << 54) - 1))
    finally:
        e64.close()
        e32.close()


def test_mixed_steps_track_fp64(sys_medium):
    """Ten full steps (one rebuild + far field, nine graph-replayed steps) in
    both modes from the same state: positions stay within the SHAKE tolerance
    (1e-4 sigma; a molecule may take one more or one fewer SHAKE sweep)."""
    from surfactant_enhanced_spreading_amd import engine
    e64, e32 = _pair(sys_medium)
    try:
        for e in (e64, e32):
            e.bootstrap()
            e.steps(10, ifar=10)
        r64, r32 = e64.state()[0], e32.state()[0]
        d = np.abs(r64 - r32).max()
        print('max position difference after 10 steps', d, flush=True)
        assert 0.0 < d <= 1e-4
    finally:
        e64.close()
        e32.close()


@pytest.mark.parametrize('nb2,p', [(3, 4), (4, 8)])
def test_mixed_m2l_bf3(sys_medium, nb2, p):
    from surfactant_enhanced_spreading_amd import engine
    es = [engine.Engine(sys_medium, nb2=nb2, p=p, precision=prec) for prec in (engine.FP64, engine.MIXED)]
    try:
        for e in es:
            e.rebuild()
            e.fmm()
        A, B = es[0].expansions(), es[1].expansions()
        for k in ('O4', 'O3', 'O2', 'L2'):
            assert np.array_equal(A[k], B[k]), k
        for k in ('L3', 'L4'):
            scale = np.abs(A[k]).max()
            err = np.abs(A[k] - B[k]).max() / scale
            print('p', p, k, 'bf16x3 error / largest coefficient', err, flush=True)
            assert 0.0 < err <= 1e-6, (k, err)
    finally:
        for e in es:
            e.close()
