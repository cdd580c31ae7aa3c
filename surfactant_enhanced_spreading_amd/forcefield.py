"""Reference force field, geometry constants and surfactant topology (reduced units).

Every value here is data that the reference fixes at compile time or in its
parameter-table setup; the tables must be identical for forces to match, so they
are restated from:

* ``src/mybox.cmns:36-43`` (sigma, epsilon, rc, delta, rs, xpi, dt),
* ``src/init_arrays.f90:91-123``  (factorial table),
* ``src/init_arrays.f90:127-225`` (42 listed intramolecular nonbonded pairs),
* ``src/init_arrays.f90:229-264`` (28 surfactant bonds),
* ``src/init_arrays.f90:266-319`` (masses),
* ``src/init_arrays.f90:322-516`` (charges, LJ parameters, combining tables),
* ``src/init_arrays.f90:520-574`` (squared constraint lengths),
* ``src/init_arrays.f90:578-681`` (angle equilibria / force constants),
* ``src/init_arrays.f90:685-840`` (dihedral coefficients),
* ``src/eval_valence.f90:13-95`` (angle site triples),
* ``src/eval_dihedral.f90:23-100`` (dihedral site quadruples).

Reduced units: length sigma = 3.166 A, energy epsilon = 650.2 J/mol, mass m_O.
Site types are numbered in one space of 32 types: 0..2 water (O, H1, H2),
3..31 surfactant sites 1..29 (``type = 2 + site``).
"""
import math

import numpy as np

# --- src/mybox.cmns:36-43 -------------------------------------------------
SIGMA = 3.166
EPSILON = 650.2
RC = 10.0 / SIGMA
DELTA = 2.0 / SIGMA
RS = RC + DELTA
# `parameter (xpi = 675.05)` is a default-real (FP32) literal: the Coulomb
# constant the reference uses is the FP32 value widened to FP64.
XPI = float(np.float32(675.05))
DT = 0.00127362
DTSQ = DT * DT
DTSQ2 = DTSQ / 2.0
P_DEFAULT = 4  # src/mybox.cmns:49

# Nose-Hoover temperature and relaxation time, src/mybox.f90:44-51
TR_DEFAULT = 3.81087
RLXTIME = 7.5 * DT

N_WATER_SITES = 3
N_SURF_SITES = 29
N_SURF_BONDS = 28
N_TYPES = 32

# --- factorials, src/init_arrays.f90:91-123 (exact double values) ---------
RIFACT = [float(math.factorial(i)) for i in range(33)]

# --- masses, src/init_arrays.f90:266-319 ----------------------------------
_m_hy = 1.00794
_m_c = 12.011
_m_o = 15.9994
_m_si = 28.0855
_m_me = _m_c + 2.0 * _m_hy
_m_methyl = _m_c + 3.0 * _m_hy
_EMS_RAW = [_m_hy, _m_o, _m_me, _m_me, _m_o, _m_me, _m_me, _m_o, _m_me, _m_me,
            _m_o, _m_me, _m_me, _m_o, _m_me, _m_me, _m_me, _m_si, _m_methyl,
            _m_o, _m_si, _m_methyl, _m_methyl, _m_methyl, _m_o, _m_si,
            _m_methyl, _m_methyl, _m_methyl]
EMS = [m / _m_o for m in _EMS_RAW]
EMW = [m / _m_o for m in (_m_o, _m_hy, _m_hy)]

# --- charges and LJ, src/init_arrays.f90:322-480 --------------------------
_sg = dict(alc_h=2.6, alc_o=3.07, opls_me=3.983, opls_o=3.047, trappe_me=3.95,
           be_si=3.385, be_o=2.955, be_methyl=3.786)
_ep = dict(alc_h=33.5, alc_o=711.76, opls_me=478.12, opls_o=817.67,
           trappe_me=382.47, be_si=2448.0, be_o=849.3, be_methyl=753.2)
_S_KIND = ['alc_h', 'alc_o', 'opls_me', 'opls_me', 'opls_o', 'opls_me', 'opls_me',
           'opls_o', 'opls_me', 'opls_me', 'opls_o', 'opls_me', 'opls_me', 'opls_o',
           'opls_me', 'trappe_me', 'trappe_me', 'be_si', 'be_methyl', 'be_o',
           'be_si', 'be_methyl', 'be_methyl', 'be_methyl', 'be_o', 'be_si',
           'be_methyl', 'be_methyl', 'be_methyl']
RQS = [0.435, -0.700, 0.265, 0.29, -0.58, 0.29, 0.29, -0.58, 0.29, 0.29, -0.58,
       0.29, 0.29, -0.58, 0.29, 0.0, 0.0, 0.3, 0.0, -0.3, 0.15, 0.0, 0.0, 0.0,
       -0.3, 0.15, 0.0, 0.0, 0.0]
SGS = [_sg[k] / SIGMA for k in _S_KIND]
EPS = [_ep[k] / EPSILON for k in _S_KIND]
RQW = [-0.8476, 0.4238, 0.4238]
SGW = [SIGMA / SIGMA, 0.0 / SIGMA, 0.0 / SIGMA]
EPW = [EPSILON / EPSILON, 0.0 / EPSILON, 0.0 / EPSILON]

# graphite carbon seen by water oxygen, src/eval_graphite.f90:10-11
SGCW6 = None  # set below (needs _pow6)
EPCW = 392.0 / EPSILON


def _pow6(x):
    # Fortran `x**6` with an integer exponent: repeated squaring
    # (x**2)**2 * x**2 is what flang emits; Python's float ** int uses the C
    # pow, which can differ by an ulp, so spell the products out.
    x2 = x * x
    return x2 * x2 * x2


SGCW6 = _pow6(3.19 / SIGMA)

# combining tables, src/init_arrays.f90:484-516
SGSS6 = [[_pow6((SGS[i] + SGS[j]) / 2.0) for j in range(29)] for i in range(29)]
EPSS = [[math.sqrt(EPS[i] * EPS[j]) for j in range(29)] for i in range(29)]
SGWS6 = [[_pow6((SGW[i] + SGS[j]) / 2.0) for j in range(29)] for i in range(3)]
EPWS = [[math.sqrt(EPW[i] * EPS[j]) for j in range(29)] for i in range(3)]
SGCS6 = [_pow6((3.19 / SIGMA + SGS[i]) / 2.0) for i in range(29)]
EPCS = [math.sqrt((392.0 / EPSILON) * EPS[i]) for i in range(29)]

# --- bonds, src/init_arrays.f90:229-264 and :520-564 ----------------------
IBOND = list(range(1, 19)) + [18, 20, 21, 21, 21, 18, 25, 26, 26, 26]
JBOND = list(range(2, 20)) + [20, 21, 22, 23, 24, 25, 26, 27, 28, 29]
_b = dict(alc_ho=0.945, alc_ome=1.430, opls_meme=1.516, opls_ome=1.410,
          trappe_meme=1.540, be_sio=1.60, be_simethyl=1.88)
_BOND_KIND = ['alc_ho', 'alc_ome', 'opls_meme', 'opls_ome', 'opls_ome', 'opls_meme',
              'opls_ome', 'opls_ome', 'opls_meme', 'opls_ome', 'opls_ome',
              'opls_meme', 'opls_ome', 'opls_ome', 'opls_meme', 'trappe_meme',
              'be_simethyl', 'be_simethyl', 'be_sio', 'be_sio', 'be_simethyl',
              'be_simethyl', 'be_simethyl', 'be_sio', 'be_sio', 'be_simethyl',
              'be_simethyl', 'be_simethyl']
DSQS = [(_b[k] * _b[k]) / (SIGMA * SIGMA) for k in _BOND_KIND]
_PI180 = 4.0 * math.atan(1.0) / 180.0
_hh = 2.0 * math.sin(109.47 * _PI180 / 2.0)
DSQW = [1.0 / (SIGMA * SIGMA), (_hh * _hh) / (SIGMA * SIGMA), 1.0 / (SIGMA * SIGMA)]

# --- angles, src/eval_valence.f90:13-95 and src/init_arrays.f90:578-681 ----
ANG_I = list(range(1, 18)) + [17, 17, 19, 19, 18, 18, 20, 20, 20, 22, 22, 23, 25,
                              25, 25, 27, 27, 28, 20]
ANG_J = list(range(2, 19)) + [18, 18, 18, 18, 25, 20, 21, 21, 21, 21, 21, 21, 26,
                              26, 26, 26, 26, 26, 18]
ANG_K = list(range(3, 20)) + [20, 25, 20, 25, 26, 21, 22, 23, 24, 24, 23, 24, 27,
                              28, 29, 28, 29, 29, 25]
_th = dict(hom=108.5, omm=108.0, mom=112.0, mmo=108.0, mmm=114.0, tet=109.5,
           sios=144.0)
_TH_KIND = ['hom', 'omm', 'mmo', 'mom', 'omm', 'mmo', 'mom', 'omm', 'mmo', 'mom',
            'omm', 'mmo', 'mom', 'omm', 'mmm'] + ['tet'] * 6 + ['sios', 'sios'] \
    + ['tet'] * 13
TH = [_th[k] * _PI180 for k in _TH_KIND]
_k = dict(a=519700.0, sios=118400.0, osio=791200.0, osime=418400.0)
_K_KIND = ['a'] * 16 + ['osime'] * 5 + ['sios', 'sios'] + ['osime'] * 6 \
    + ['osime'] * 3 + ['osime'] * 3 + ['osio']
RKTH = [_k[k] / EPSILON for k in _K_KIND]

# --- dihedrals, src/eval_dihedral.f90:23-100, src/init_arrays.f90:685-840 --
DIH_I = list(range(1, 17)) + [16, 16, 17, 17, 19, 19, 18, 18, 18, 18, 18, 18, 20, 21]
DIH_J = list(range(2, 18)) + [17, 17, 18, 18, 18, 18, 20, 20, 20, 25, 25, 25, 18, 20]
DIH_K = list(range(3, 19)) + [18, 18, 20, 25, 20, 25, 21, 21, 21, 26, 26, 26, 25, 18]
DIH_L = list(range(4, 20)) + [20, 25, 21, 26, 21, 26, 22, 23, 24, 27, 28, 29, 26, 25]
_dv = dict(alcohol=(3490.0, -486.0, 3130.0), oxmmox=(2940.0, -888.0, 12810.0),
           mmoxm=(19860.0, -5853.0, 8956.0), moxmm=(19860.0, -5853.0, 8956.0),
           oxmmm=(2940.0, -888.0, 12810.0), trappe=(5903.8, -1134.0, 13159.0))
_DV_KIND = ['alcohol', 'oxmmox', 'mmoxm', 'moxmm', 'oxmmox', 'mmoxm', 'moxmm',
            'oxmmox', 'mmoxm', 'moxmm', 'oxmmox', 'mmoxm', 'moxmm', 'oxmmm',
            'trappe', 'trappe', 'trappe', 'trappe']
DV_BE = 3770.0
DVA = [_dv[k][0] / 2.0 / EPSILON for k in _DV_KIND] + [DV_BE / EPSILON] * 12
DVB = [_dv[k][1] / 2.0 / EPSILON for k in _DV_KIND] + [0.0 / EPSILON] * 12
DVC = [_dv[k][2] / 2.0 / EPSILON for k in _DV_KIND] + [0.0 / EPSILON] * 12
N_DIH_OPLS = 18  # dihedrals 1..18: OPLS form; 19..30: Berendsen cos(3 phi)

# --- listed intramolecular nonbonded pairs, src/init_arrays.f90:127-225 ----
ISELF = [16] * 8 + [17] * 6 + [19] * 6 + [20] * 3 + [21] * 4 + [22] * 5 + [23] * 5 \
    + [24] * 5
JSELF = [21, 22, 23, 24, 26, 27, 28, 29, 22, 23, 24, 27, 28, 29, 22, 23, 24, 27, 28,
         29, 27, 28, 29, 26, 27, 28, 29, 25, 26, 27, 28, 29, 25, 26, 27, 28, 29, 25,
         26, 27, 28, 29]
# plus every pair (j, k) with j <= 15 and k >= j + 4, src/eval_surfactant_internal.f90:155-169


def internal_pairs():
    """1-based (j, k) surfactant site pairs that get intramolecular LJ + Coulomb,
    in the order ``eval_surfactant_internal`` visits them
    (``src/eval_surfactant_internal.f90:155-244``)."""
    pairs = [(j, k) for j in range(1, 16) for k in range(j + 4, 30)]
    pairs += list(zip(ISELF, JSELF))
    return pairs


def dihedral_clamp(i):
    """cos(phi) clamp used for dihedral i (1-based): OPLS terms clamp to
    +-0.99999999999 (``src/eval_dihedral.f90:142-143``), Berendsen terms to
    +-0.999999 (``src/eval_dihedral_berendsen.f90:142-143``)."""
    return 0.99999999999 if i <= N_DIH_OPLS else 0.999999


def type_tables():
    """Tables over the 32-type space used by the kernels.

    Returns a dict of numpy arrays:
      q[32]            charge (no xpi factor),
      mass[32]         mass / m_O,
      lj_c[32,32]      (48*eps_ij)*s6_ij   (the prefactor of r^-6 in the pair force),
      lj_s6[32,32]     s6_ij = sigma_ij^6,
      gr_c[32], gr_s6[32]  the same for carbon (graphite) partners.
    A zero lj_c entry means the pair has no LJ term.  The water O-O entry is
    (48*1)*1 with s6 = 1, i.e. the reference's `48 r6 (r6 - 1/2)` exactly.
    """
    q = np.zeros(N_TYPES)
    mass = np.zeros(N_TYPES)
    lj_c = np.zeros((N_TYPES, N_TYPES))
    lj_s6 = np.zeros((N_TYPES, N_TYPES))
    gr_c = np.zeros(N_TYPES)
    gr_s6 = np.zeros(N_TYPES)
    q[0:3] = RQW
    mass[0:3] = EMW
    q[3:] = RQS
    mass[3:] = EMS
    # water O - water O
    lj_c[0, 0] = 48.0 * 1.0 * 1.0
    lj_s6[0, 0] = 1.0
    for n in range(29):
        # water O - surfactant n (src/evalw1s.f90: epws(1,n), sgws6(1,n))
        c = 48.0 * EPWS[0][n] * SGWS6[0][n]
        lj_c[0, 3 + n] = lj_c[3 + n, 0] = c
        lj_s6[0, 3 + n] = lj_s6[3 + n, 0] = SGWS6[0][n]
        for m in range(29):
            lj_c[3 + n, 3 + m] = 48.0 * EPSS[n][m] * SGSS6[n][m]
            lj_s6[3 + n, 3 + m] = SGSS6[n][m]
        gr_c[3 + n] = 48.0 * EPCS[n] * SGCS6[n]
        gr_s6[3 + n] = SGCS6[n]
    gr_c[0] = 48.0 * EPCW * SGCW6
    gr_s6[0] = SGCW6
    return dict(q=q, mass=mass, lj_c=lj_c, lj_s6=lj_s6, gr_c=gr_c, gr_s6=gr_s6)
