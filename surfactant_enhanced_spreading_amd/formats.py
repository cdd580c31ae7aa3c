"""The reference's file formats, read and written the way its routines do.

* ``data_cont2.dat`` -- the starting configuration ``read_config`` reads
  (``src/read_config.f90:19-34``): for every site, waters first as (O, H1, H2)
  then surfactants 29 sites each, one position line and one velocity line, both
  ``(f8.3,f9.3,f9.3)`` in reduced units.  On reading, ``exp(-9)`` is added to
  every coordinate and a velocity field ``w`` becomes ``(w - 0.5) * 0.01`` (the
  v*dt storage convention).
* ``graphite.dat`` -- the carbon sheet ``make_graphite`` writes and
  ``read_graphite`` reads, ``(f10.4,f11.4,f11.4)`` per atom
  (``src/make_graphite.f90:125-129``, ``src/read_graphite.f90:19-22``);
  ``make_graphite()`` restates the generator itself (22,956 atoms).
* PDB snapshots as ``write_pdb`` writes them (``src/write_pdb.f90:19-169``):
  waters in files of at most 33,333 molecules (``dataNNNNNNNa.pdb``,
  ``...b.pdb``, ...), surfactants in ``dataNNNNNNNs.pdb``, coordinates in
  Angstrom.

Fortran fixed-format fields are emulated exactly: an F input field without a
decimal point takes the implied decimals, blanks are ignored, and output
fields are right-justified, filled with ``*`` when the value does not fit.
"""
import math
import os

import numpy as np

from . import forcefield as ff
from .builder import System

CONFIG_SHIFT = math.exp(-9.0)   # src/read_config.f90:28-30
PDB_WATERS_PER_FILE = 33333     # src/write_pdb.f90:21-23 (ilimit)
PDB_FORMAT = (('a', 6), ('i', 5), ('a', 5), ('a', 4), ('i', 6), ('f', 12, 3), ('f', 8, 3),
              ('f', 8, 3))      # '(a6,i5,a5,a4,i6,f12.3,f8.3,f8.3)'
# element of each surfactant site in the PDB file, src/write_pdb.f90:106-164
SURF_ELEMENTS = ('H', 'O', 'C', 'C', 'O', 'C', 'C', 'O', 'C', 'C', 'O', 'C', 'C', 'O', 'C', 'C',
                 'C', 'Si', 'C', 'O', 'Si', 'C', 'C', 'C', 'O', 'Si', 'C', 'C', 'C')


# ---------------------------------------------------------------- fields

def f_field(x, w, d):
    """Fw.d output: right-justified, '*' * w on overflow."""
    s = '%.*f' % (d, x)
    if s.startswith('0.') and len(s) > w:
        s = s[1:]                  # the optional leading zero goes first
    elif s.startswith('-0.') and len(s) > w:
        s = '-' + s[2:]
    return s.rjust(w) if len(s) <= w else '*' * w


def i_field(n, w):
    s = str(int(n))
    return s.rjust(w) if len(s) <= w else '*' * w


def a_field(text, w):
    """Aw output of a shorter string: right-justified (leading blanks)."""
    return text.rjust(w) if len(text) <= w else text[:w]


def parse_f(field, d):
    """Fw.d input: blanks ignored; without a decimal point or exponent the last
    d digits are decimals; an empty field reads as zero."""
    s = field.replace(' ', '')
    if not s:
        return 0.0
    if '.' in s or any(c in s for c in 'eEdD'):
        return float(s.replace('d', 'e').replace('D', 'e'))
    return int(s) / 10.0 ** d


def _split(line, widths):
    out, p = [], 0
    for w in widths:
        out.append(line[p:p + w])
        p += w
    return out


# ---------------------------------------------------------------- data_cont2.dat

def read_config(path, n_water, n_surf):
    """Positions and velocities of a data_cont2.dat file (src/read_config.f90:19-34)
    as a System (sigma units; velocities as v*dt)."""
    n = 3 * n_water + 29 * n_surf
    pos = np.empty((n, 3))
    vel = np.empty((n, 3))
    with open(path) as fh:
        for i in range(n):
            p = fh.readline()
            v = fh.readline()
            if not v:
                raise ValueError('%s: %d sites expected, file ends at site %d' % (path, n, i))
            pos[i] = [parse_f(f, 3) for f in _split(p.rstrip('\n'), (8, 9, 9))]
            vel[i] = [parse_f(f, 3) for f in _split(v.rstrip('\n'), (8, 9, 9))]
    pos += CONFIG_SHIFT
    vel = (vel - 0.5) * 0.01
    return System(n_water, n_surf, pos, vel)


def write_config(path, sys_):
    """The inverse of read_config: positions as written, velocities encoded as
    v / 0.01 + 0.5 (each to the 1e-3 the format keeps)."""
    w = sys_.vel / 0.01 + 0.5
    with open(path, 'w') as fh:
        for i in range(sys_.n_sites):
            for row in (sys_.pos[i], w[i]):
                fh.write(f_field(row[0], 8, 3) + f_field(row[1], 9, 3) + f_field(row[2], 9, 3) + '\n')


# ---------------------------------------------------------------- graphite.dat

def make_graphite():
    """The reference's graphite generator, src/make_graphite.f90:17-112: two
    layers of a hexagonal sheet (ix = 81 columns), the second shifted by one
    C-C bond and 3.41 A down, centred in x and y, top layer at z = -16 sigma.
    Returns (22956, 3) float64 in sigma units, in the file's order."""
    pi = 4.0 * math.atan(1.0)
    sigma = 3.166
    ccbond = 1.421 / sigma
    xshift = math.cos(60.0 * pi / 180.0) * ccbond
    yshift = math.sin(60.0 * pi / 180.0) * ccbond
    ix = 81
    iy = int(math.sin(60.0 * pi / 180.0) * ix)
    ins = 2 * ix * iy + 2 * (iy - 1)
    xcarbon, ycarbon, zcarbon = 0.0, 0.0, -16.0
    x, y, z = [], [], []
    for i in range(1, ix + 1):
        for j in range(1, iy + 1):
            yup = yshift if i % 2 == 0 else 0.0
            xa = i * (ccbond + xshift)
            ya = j * 2.0 * yshift + yup
            x += [xa, xa + ccbond]
            y += [ya, ya]
            z += [zcarbon, zcarbon]
    for j in range(1, iy):
        x.append(ccbond)
        y.append(j * 2.0 * yshift + yshift)
        z.append(zcarbon)
    for j in range(1, iy):
        x.append((ix + 1) * (ccbond + xshift))
        y.append((j + 1) * 2.0 * yshift - yshift)
        z.append(zcarbon)
    for j in range(ins):
        x.append(x[j] + ccbond)
        y.append(y[j])
        z.append(z[j] - 3.41 / sigma)
    x, y, z = np.array(x), np.array(y), np.array(z)
    cx = (x.max() + x.min()) / 2.0
    cy = (y.max() + y.min()) / 2.0
    x = x - cx + xcarbon
    y = y - cy + ycarbon
    return np.stack([x, y, z], axis=1)


def write_graphite(path, g):
    with open(path, 'w') as fh:
        for a in g:
            fh.write(f_field(a[0], 10, 4) + f_field(a[1], 11, 4) + f_field(a[2], 11, 4) + '\n')


def write_graphite_pdb(path, g):
    """graphite.pdb as make_graphite writes it beside graphite.dat
    (src/make_graphite.f90:131-135): every carbon as residue GPH, in Angstrom."""
    with open(path, 'w') as fh:
        for j, a in enumerate(g):
            fh.write(_pdb_line(j + 1, 'C', 'GPH', j + 1, a * 3.166))


def read_graphite(path, count=None):
    """Carbon positions of a graphite.dat file (the reference reads the first
    22956 lines, src/read_graphite.f90:20-22; count=None reads them all)."""
    out = []
    with open(path) as fh:
        for k, line in enumerate(fh):
            if count is not None and k >= count:
                break
            out.append([parse_f(f, 4) for f in _split(line.rstrip('\n'), (10, 11, 11))])
    return np.array(out, dtype=np.float64).reshape(-1, 3)


def load_system(config_path, n_water, n_surf, graphite_path=None):
    """A System from the reference's two input files."""
    s = read_config(config_path, n_water, n_surf)
    if graphite_path is not None:
        s.graphite = read_graphite(graphite_path)
    return s


# ---------------------------------------------------------------- PDB

def _pdb_line(serial, element, resname, resid, xyz):
    vals = ('HETATM', serial, element, resname, resid, xyz[0], xyz[1], xyz[2])
    out = []
    for spec, v in zip(PDB_FORMAT, vals):
        if spec[0] == 'a':
            out.append(a_field(v, spec[1]))
        elif spec[0] == 'i':
            out.append(i_field(v, spec[1]))
        else:
            out.append(f_field(v, spec[1], spec[2]))
    return ''.join(out) + '\n'


def pdb_names(icq, n_water):
    """File names write_pdb uses for snapshot icq (src/write_pdb.f90:25-73):
    the water files, then the surfactant file."""
    stem = 'data' + ('%07d' % icq if icq < 10 ** 7 else str(icq)[:7])
    nfiles = int(n_water / (PDB_WATERS_PER_FILE + 0.0001)) + 1
    return [stem + 'abcdefghij'[k] + '.pdb' for k in range(nfiles)], stem + 's.pdb'


def write_pdb(directory, sys_, icq):
    """Write snapshot icq of sys_ as the reference's write_pdb would: returns the
    paths written.  Coordinates are sigma units times sigma (Angstrom)."""
    sigma = ff.SIGMA
    wnames, sname = pdb_names(icq, sys_.n_water)
    paths = []
    nw = sys_.n_water
    for k, name in enumerate(wnames):
        lo = k * PDB_WATERS_PER_FILE
        hi = min(nw, lo + PDB_WATERS_PER_FILE)
        p = os.path.join(directory, name)
        with open(p, 'w') as fh:
            for i in range(hi - lo):
                j = 3 * (lo + i)
                for a, el in enumerate(('O', 'H', 'H')):
                    fh.write(_pdb_line(3 * i + a + 1, el, 'WAT', i + 1, sys_.pos[j + a] * sigma))
        paths.append(p)
    p = os.path.join(directory, sname)
    base = 3 * nw
    with open(p, 'w') as fh:
        serial = 0
        for m in range(sys_.n_surf):
            for k in range(29):
                serial += 1
                fh.write(_pdb_line(serial, SURF_ELEMENTS[k], 'SRF', m + 1,
                                   sys_.pos[base + 29 * m + k] * sigma))
    paths.append(p)
    return paths
