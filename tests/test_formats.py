"""The reference's file formats (data_cont2.dat, graphite.dat, PDB snapshots):
against digests of the reference's own output (tests/golden/formats.json,
written by the compiled Fortran through tests/golden/make_goldens.py), and
live against the compiled routines where the reference tree is present."""
import hashlib
import json
import os
import subprocess

import numpy as np
import pytest

from conftest import file_format_system
from surfactant_enhanced_spreading_amd import formats as fmt

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLD = json.load(open(os.path.join(ROOT, 'tests', 'golden', 'formats.json')))
REF_LIB = os.path.join(ROOT, 'oracle', '_ref', 'libmybox_ref.so')
needs_ref = pytest.mark.skipif(not os.path.exists(REF_LIB),
                               reason='reference Fortran not built here (make -C oracle ref)')


def sha(path):
    with open(path, 'rb') as fh:
        return hashlib.sha256(fh.read()).hexdigest()


@pytest.fixture(scope='module')
def ffs():
    return file_format_system()


def test_fixed_fields():
    # F input: implied decimals without a point, blanks ignored, empty = 0
    assert fmt.parse_f('   1234', 3) == 1.234
    assert fmt.parse_f(' -12.5  ', 3) == -12.5
    assert fmt.parse_f('        ', 3) == 0.0
    assert fmt.parse_f(' 1.5d2', 3) == 150.0
    # F/I/A output: right-justified, stars on overflow
    assert fmt.f_field(-3.14159, 8, 3) == '  -3.142'
    assert fmt.f_field(123456.0, 8, 3) == '********'
    assert fmt.i_field(7, 5) == '    7' and fmt.i_field(123456, 5) == '*****'
    assert fmt.a_field('Si', 5) == '   Si'


def test_graphite_generator_matches_reference_file(tmp_path):
    g = fmt.make_graphite()
    assert g.shape == (22956, 3)
    p = tmp_path / 'graphite.dat'
    fmt.write_graphite(str(p), g)
    assert sha(str(p)) == GOLD['graphite']['sha256']
    lines = p.read_text().split('\n')
    assert lines[0] == GOLD['graphite']['first'] and lines[-2] == GOLD['graphite']['last']
    q = tmp_path / 'graphite.pdb'
    fmt.write_graphite_pdb(str(q), g)
    assert sha(str(q)) == GOLD['pdb']['graphite.pdb']
    # the reader gives back the file's values (4 decimals)
    back = fmt.read_graphite(str(p))
    assert np.abs(back - g).max() <= 0.5e-4 + 1e-12


def test_pdb_matches_reference(tmp_path, ffs):
    paths = fmt.write_pdb(str(tmp_path), ffs, 42)
    assert sorted(os.path.basename(p) for p in paths) == ['data0000042a.pdb', 'data0000042s.pdb']
    for p in paths:
        assert sha(p) == GOLD['pdb'][os.path.basename(p)], p


def test_pdb_file_split():
    # 33,333 waters per file: names a, b, c ... then s (src/write_pdb.f90:71-101)
    w, s = fmt.pdb_names(7, 70000)
    assert w == ['data0000007a.pdb', 'data0000007b.pdb', 'data0000007c.pdb']
    assert s == 'data0000007s.pdb'
    assert fmt.pdb_names(1234567, 33333)[0] == ['data1234567a.pdb']


def test_config_reader_matches_reference(tmp_path, ffs):
    p = tmp_path / 'data_cont2.dat'
    fmt.write_config(str(p), ffs)
    assert sha(str(p)) == GOLD['data_cont2_sha256']
    s = fmt.read_config(str(p), ffs.n_water, ffs.n_surf)
    rows = np.concatenate([s.pos, s.vel], axis=1)
    rows = rows[np.lexsort(rows.T[::-1])]
    assert len(rows) == GOLD['read_config_rows']
    assert hashlib.sha256(np.ascontiguousarray(rows, dtype='<f8').tobytes()).hexdigest() == \
        GOLD['read_config_sha256']


def test_config_round_trip(tmp_path, ffs):
    p = tmp_path / 'data_cont2.dat'
    fmt.write_config(str(p), ffs)
    s = fmt.read_config(str(p), ffs.n_water, ffs.n_surf)
    # positions keep 3 decimals and gain exp(-9); velocities keep 1e-5 of v dt
    assert np.abs(s.pos - (ffs.pos + fmt.CONFIG_SHIFT)).max() <= 0.5e-3 + 1e-9
    assert np.abs(s.vel - ffs.vel).max() <= 0.5e-5 + 1e-12
    # writing what was read (less the shift) and reading again is exact
    # (the text may differ only in the sign of a zero field)
    q = tmp_path / 'again.dat'
    s2 = s.copy()
    s2.pos = s2.pos - fmt.CONFIG_SHIFT
    fmt.write_config(str(q), s2)
    s3 = fmt.read_config(str(q), ffs.n_water, ffs.n_surf)
    assert np.array_equal(s3.pos, s.pos) and np.array_equal(s3.vel, s.vel)


def test_config_reader_rejects_short_file(tmp_path):
    p = tmp_path / 'short.dat'
    p.write_text('   1.000    2.000    3.000\n   0.500    0.500    0.500\n')
    with pytest.raises(ValueError):
        fmt.read_config(str(p), 1, 0)


@needs_ref
def test_live_make_graphite(tmp_path):
    import refharness as rh
    subprocess.run([rh.MAKE_GRAPHITE], cwd=str(tmp_path), check=True, capture_output=True)
    ours = tmp_path / 'ours.dat'
    fmt.write_graphite(str(ours), fmt.make_graphite())
    assert (tmp_path / 'graphite.dat').read_text() == ours.read_text()


@needs_ref
def test_live_read_config_and_pdb(tmp_path, ffs):
    import refharness as rh
    ref = rh.RefLib()
    fmt.write_config(str(tmp_path / 'data_cont2.dat'), ffs)
    s = fmt.read_config(str(tmp_path / 'data_cont2.dat'), ffs.n_water, ffs.n_surf)
    em = rh.Emulator(ref, s)
    assert em.read_config_mismatches(str(tmp_path)) == []
    d1, d2 = tmp_path / 'ref', tmp_path / 'ours'
    d1.mkdir()
    d2.mkdir()
    ref.write_pdb(ffs, 42, str(d1))
    for p in fmt.write_pdb(str(d2), ffs, 42):
        assert open(p).read() == open(str(d1 / os.path.basename(p))).read()
