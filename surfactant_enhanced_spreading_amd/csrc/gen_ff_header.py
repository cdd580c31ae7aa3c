"""Write ff_reference.h: the reference force field as a C initializer.

The Fortran drop-in library (dropin.cpp) takes the per-type numbers from the
reference's own COMMON tables at run time; what the reference hard-codes in
its routines instead (angle and dihedral site lists, src/eval_valence.f90:13-95
and src/eval_dihedral.f90:23-100; the intramolecular pair rule of
src/eval_surfactant_internal.f90:12-110; the water-carbon LJ constants of
src/eval_graphite.f90; the dihedral clamps; the PARAMETERs of
src/mybox.cmns:36-49) comes from this header, generated from forcefield.py so
that both sides use one table.

    python -m surfactant_enhanced_spreading_amd.csrc.gen_ff_header
"""
import os

from surfactant_enhanced_spreading_amd.engine import reference_forcefield

HERE = os.path.dirname(os.path.abspath(__file__))
OUT = os.path.join(HERE, 'ff_reference.h')


def _arr(vals, fmt):
    items = [fmt(v) for v in vals]
    lines, cur = [], ''
    for it in items:
        if len(cur) + len(it) + 2 > 96:
            lines.append(cur)
            cur = ''
        cur += it + ', '
    lines.append(cur)
    return '{\n        ' + '\n        '.join(l.rstrip() for l in lines) + '\n    }'


def render():
    ff = reference_forcefield()
    dbl = lambda v: repr(float(v))
    out = ['/* Generated by gen_ff_header.py from forcefield.py -- do not edit. */',
           '#ifndef SESMD_FF_REFERENCE_H', '#define SESMD_FF_REFERENCE_H',
           '#include <string.h>', '#include "sesmd.h"', '',
           'static inline void sesmd_reference_forcefield(sesmd_forcefield *f)', '{']
    for name, typ, *_ in ff._fields_:
        val = getattr(ff, name)
        if hasattr(val, '__len__'):
            ctype = 'int' if 'c_int' in repr(typ) else 'double'
            fmt = (lambda v: str(int(v))) if ctype == 'int' else dbl
            out.append('    static const %s %s_[%d] = %s;' % (ctype, name, len(val),
                                                            _arr(list(val), fmt)))
            out.append('    memcpy(f->%s, %s_, sizeof %s_);' % (name, name, name))
        else:
            out.append('    f->%s = %s;' % (name, dbl(val) if isinstance(val, float)
                                           else str(val)))
    out += ['}', '', '#endif', '']
    return '\n'.join(out)


def write():
    text = render()
    old = open(OUT).read() if os.path.exists(OUT) else None
    if old != text:
        with open(OUT, 'w') as fh:
            fh.write(text)
    return OUT


if __name__ == '__main__':
    print(write())
