"""The force-field tables restated in forcefield.py equal, bit for bit, the
COMMON tables the reference's init_arrays fills (tests/golden/tables.npz)."""
import os

import numpy as np

from surfactant_enhanced_spreading_amd import forcefield as ff

G = os.path.join(os.path.dirname(os.path.abspath(__file__)), 'golden', 'tables.npz')


def test_tables_bit_exact():
    z = np.load(G, allow_pickle=False)
    pairs = {'emw': ff.EMW, 'ems': ff.EMS, 'sgw': ff.SGW, 'sgs': ff.SGS, 'epw': ff.EPW,
             'eps': ff.EPS, 'rqw': ff.RQW, 'rqs': ff.RQS, 'dsqw': ff.DSQW, 'dsqs': ff.DSQS,
             'th': ff.TH, 'rkth': ff.RKTH, 'dva': ff.DVA, 'dvb': ff.DVB, 'dvc': ff.DVC,
             'epss': ff.EPSS, 'epws': ff.EPWS, 'sgcs6': ff.SGCS6, 'epcs': ff.EPCS,
             'rifact': ff.RIFACT, 'iself': ff.ISELF, 'jself': ff.JSELF, 'ibond': ff.IBOND,
             'jbond': ff.JBOND}
    for k, v in pairs.items():
        assert np.array_equal(np.asarray(v, dtype=z[k].dtype), z[k]), k
    # sigma^6 tables: Fortran x**6 and the restated product may differ in the last bit
    for k, v in (('sgss6', ff.SGSS6), ('sgws6', ff.SGWS6)):
        assert np.abs(np.asarray(v) - z[k]).max() <= 4e-16 * np.abs(z[k]).max(), k


def test_type_tables_consistent():
    T = ff.type_tables()
    assert T['lj_c'][0, 0] == 48.0 and T['lj_s6'][0, 0] == 1.0
    assert np.all(T['lj_c'][1:3, :] == 0.0) and np.all(T['lj_c'][:, 1:3] == 0.0)
    assert np.allclose(T['lj_c'], T['lj_c'].T, rtol=0, atol=0)
    assert abs(sum(ff.RQS)) < 1e-12 and abs(sum(ff.RQW)) < 1e-12
    assert ff.XPI == 675.0499877929688
