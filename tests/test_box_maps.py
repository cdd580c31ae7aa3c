"""The box-map functions of libsesmd_fortran.so (src/ibox.f90:1-220) against
the reference's own compiled ibox.f90, bit for bit.

The reference library is loaded first (RTLD_GLOBAL): its COMMON blocks hold the
geometry both sets of functions read (/cmn01/-/cmn04/, /cmn17/), filled here as
src/init_box_dimensions.f90 fills them for the 27-rank layout.  Inputs: every
index of the integer maps; random positions over and beyond the global cube;
and the edge cases int() truncation and the half-open tests make sharp --
points exactly on box faces, one ulp either side, and small negative offsets
where truncation toward zero and floor differ.

Needs oracle/_ref/libmybox_ref.so (make -C oracle ref); skipped where absent.
No GPU is used: these are host functions.
"""
import ctypes
import os

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF_SO = os.path.join(ROOT, 'oracle', '_ref', 'libmybox_ref.so')
SHIM_SO = os.path.join(ROOT, 'surfactant_enhanced_spreading_amd', 'libsesmd_fortran.so')

pytestmark = pytest.mark.skipif(
    not (os.path.exists(REF_SO) and os.path.exists(SHIM_SO)),
    reason='reference Fortran (make -C oracle ref) or libsesmd_fortran.so not built')

POS3 = ('ibox2', 'ibox3', 'ibox4', 'ibox64', 'ibox1000')
POS6 = ('ibox64mod', 'isnearest', 'iscentral')
INT1 = {'ib64to216': 64, 'ib64to512': 64, 'ib64to1000': 64, 'ib8to216': 8, 'ib512to64': 512}


@pytest.fixture(scope='module')
def libs():
    import refharness as rh
    ref = rh.RefLib()
    shim = ctypes.CDLL(SHIM_SO)
    dp, ip = ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_int)
    for L in (ref.lib, shim):
        for n in POS3:
            getattr(L, n + '_').restype = ctypes.c_int
            getattr(L, n + '_').argtypes = [dp] * 3
        for n in POS6:
            getattr(L, n + '_').restype = ctypes.c_int
            getattr(L, n + '_').argtypes = [dp] * 6
        for n in INT1:
            getattr(L, n + '_').restype = ctypes.c_int
            getattr(L, n + '_').argtypes = [ip]
    rs = 10.0 / 3.166 + 2.0 / 3.166
    rsd = 12 * rs
    ref['ibdx3'][0] = ref['ibdy3'][0] = ref['ibdz3'][0] = 6
    ref['ibdx4'][0] = ref['ibdy4'][0] = ref['ibdz4'][0] = 12
    ref['rsdx'][0] = ref['rsdy'][0] = ref['rsdz'][0] = rsd
    ref['rhsx'][0] = ref['rhsy'][0] = ref['rhsz'][0] = rsd / 2.0
    return ref, shim, rs


def _d(v):
    return ctypes.byref(ctypes.c_double(float(v)))


def _edges(c, rs):
    """Faces of the boxes around centre c along one axis at multiples of rs/2,
    one ulp either side, and small offsets."""
    out = []
    for k in range(-12, 13):
        x = c + 0.5 * k * rs
        out += [x, np.nextafter(x, -np.inf), np.nextafter(x, np.inf), x + 1e-13, x - 1e-13]
    return out


def test_integer_maps(libs):
    ref, shim, _ = libs
    for n, top in INT1.items():
        for i in range(1, top + 1):
            a = getattr(ref.lib, n + '_')(ctypes.byref(ctypes.c_int(i)))
            b = getattr(shim, n + '_')(ctypes.byref(ctypes.c_int(i)))
            assert a == b, (n, i, a, b)


@pytest.mark.parametrize('rank', [(0, 0, 0), (1, 1, 1), (2, 0, 1)])
def test_position_maps(libs, rank):
    ref, shim, rs = libs
    rhs = ref['rhsx'][0]
    cen = [(4 * rs) * (i + 1) - rhs - 2 * rs for i in rank]  # rcen2 of the rank
    ref['rcen2xtmp'][0], ref['rcen2ytmp'][0], ref['rcen2ztmp'][0] = cen
    rng = np.random.default_rng(sum(rank) + 17)
    pts = list(rng.uniform(-1.2 * rhs, 1.2 * rhs, size=(400, 3)))
    for a in range(3):
        for x in _edges(cen[a], rs):
            p = list(rng.uniform(-rhs, rhs, size=3))
            p[a] = x
            pts.append(p)
    # global-grid faces (ibox2/3/4): x + rhs a multiple of rs, and near -rhs
    for k in range(0, 13):
        x = -rhs + k * rs
        for v in (x, np.nextafter(x, -np.inf), np.nextafter(x, np.inf), -rhs - 0.3 * rs):
            pts.append([v, v, -v])
    nfun = 0
    for p in pts:
        for n in POS3:
            a = getattr(ref.lib, n + '_')(*[_d(v) for v in p])
            b = getattr(shim, n + '_')(*[_d(v) for v in p])
            assert a == b, (n, p, a, b)
            nfun += 1
        for c in (cen, [cen[0] + rs, cen[1] - rs, cen[2]], p):
            q = list(p) + list(c)
            for n in POS6:
                a = getattr(ref.lib, n + '_')(*[_d(v) for v in q])
                b = getattr(shim, n + '_')(*[_d(v) for v in q])
                assert a == b, (n, q, a, b)
                nfun += 1
    assert nfun > 5000
    # the truncation toward zero is exercised: some points sit just below a
    # grid's first box (int() = 0 where floor would give -1)
    low = [cen[0] - 2 * rs - 0.3 * rs, cen[1], cen[2]]
    assert shim.ibox64_(*[_d(v) for v in low]) == ref.lib.ibox64_(*[_d(v) for v in low])


def test_half_open_faces(libs):
    """isCentral / isnearest include the low face and exclude the high one."""
    _, shim, rs = libs
    c = [0.3, -1.1, 2.0]
    args = lambda p: [_d(v) for v in list(p) + c]
    lo = [c[0] - rs / 2.0, c[1], c[2]]
    hi = [c[0] + rs / 2.0, c[1], c[2]]
    assert shim.iscentral_(*args(lo)) == 1
    assert shim.iscentral_(*args(hi)) == 0
    lo = [c[0], c[1] - 1.5 * rs, c[2]]
    hi = [c[0], c[1], c[2] + 1.5 * rs]
    assert shim.isnearest_(*args(lo)) == 1
    assert shim.isnearest_(*args(hi)) == 0
