"""Write the golden fixtures from the reference's own compiled Fortran routines.

Run in a checkout that has the reference tree and the compiled subset
(`make -C oracle ref`):

    python tests/golden/make_goldens.py

Outputs (numpy .npz, compressed), all values produced by the reference code
through oracle/refharness.py (the 27-rank COMMON emulation):
  tables.npz        the COMMON parameter tables init_arrays fills
  ylm.npz           Ylm / rLegendreP on a grid of arguments
  small_system.npz  inputs + bootstrap forces per component, per reference
                    kernel, classes, isub, far-field expansions, one full step
  medium_total.npz  inputs + total bootstrap force of a larger system
  r2_system.npz     the reference's workload of record (9,997 waters + 475
                    surfactants, its PARAMETER sizes): inputs as the fields of
                    the data_cont2.dat it reads (integers in units of 1e-3;
                    formats.read_config's shift and velocity encoding apply),
                    bootstrap total
                    force and one full step (part A, rebuild, forces, part B)
                    at every 7th site
  velocities.json   the first 64 values of the reference's generator
                    (src/random.f90) from the start-up seed, as hex floats
  formats.json      SHA-256 digests of the reference's own file output
                    (make_graphite's graphite.dat, write_pdb's snapshot of the
                    file-format test system) and of the positions/velocities
                    read_config stores from that system's data_cont2.dat
"""
import hashlib
import json
import subprocess
import tempfile
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, 'oracle'))
sys.path.insert(0, os.path.join(ROOT, 'tests'))

import refharness as rh  # noqa: E402
from conftest import small_system, medium_system, file_format_system  # noqa: E402
from surfactant_enhanced_spreading_amd import formats as fmt  # noqa: E402


def tables(ref):
    keys = ['emw', 'ems', 'sgw', 'sgs', 'epw', 'eps', 'rqw', 'rqs', 'dsqw', 'dsqs', 'th',
            'rkth', 'dva', 'dvb', 'dvc', 'sgww6', 'epww', 'sgss6', 'epss', 'sgws6', 'epws',
            'sgcs6', 'epcs', 'rifact', 'iself', 'jself', 'ibond', 'jbond']
    np.savez_compressed(os.path.join(HERE, 'tables.npz'),
                        **{k: np.array(ref[k]) for k in keys})


def ylm(ref):
    rng = np.random.default_rng(11)
    ang = rng.uniform([0.0, -np.pi], [np.pi, np.pi], size=(12, 2))
    ang[0] = (0.3, 0.0)
    ang[1] = (np.pi / 2, np.pi / 3)
    rows = []
    vals = []
    for a, b in ang:
        for n in range(0, 9):
            for m in range(-n, n + 1):
                rows.append((n, m, a, b))
                vals.append(ref.ylm(n, m, a, b))
    xs = np.concatenate([np.linspace(-1, 1, 9), rng.uniform(-1, 1, 8)])
    leg = [(l, m, x, ref.legendre(l, m, x)) for x in xs for l in range(-1, 9) for m in range(0, 9)
           if not (l == -1 and m > 0)]
    np.savez_compressed(os.path.join(HERE, 'ylm.npz'), args=np.array(rows), vals=np.array(vals),
                        leg=np.array(leg))


def small(ref):
    s = small_system()
    em = rh.Emulator(ref, s)
    out = dict(n_water=s.n_water, n_surf=s.n_surf, pos=s.pos, vel=s.vel, graphite=s.graphite)
    out['cls'] = em.cls
    out['anc'] = em.anc[:, 0] + 12 * (em.anc[:, 1] + 12 * em.anc[:, 2])
    out['isub'] = em.isub_masks()
    out['F_near'] = em.forces(far=False, intra=False, wall=False, graphite=False)
    out['F_far'] = em.forces(near=False, intra=False, wall=False, graphite=False)
    out['O4'] = em.O4.reshape(-1, rh.LP)
    L4 = np.zeros((12, 12, 12, rh.LP), dtype=np.complex128)
    for r, rk in enumerate(em.ranks()):
        for i in range(64):
            g = em._glob(rk, i)
            L4[g[2], g[1], g[0]] = em.cL4[r][:, i]
    L4 = L4.reshape(-1, rh.LP)
    occ = np.nonzero(np.abs(out['O4']).max(axis=1))[0].astype(np.int32)
    out['occ'] = occ  # boxes holding receivers: O4 and L4 are kept there
    out['O4'] = out['O4'][occ]
    out['L4'] = L4[occ]
    out['O3'] = em.O3.reshape(-1, rh.LP)
    out['O2'] = em.O2
    out['F_intra'] = em.forces(near=False, far=False, wall=False, graphite=False)
    out['F_wall'] = em.forces(near=False, far=False, intra=False, graphite=False)
    out['F_graphite'] = em.forces(near=False, far=False, intra=False, wall=False)
    out['F'] = em.forces()
    names = []
    for k in rh.NEAR_KERNELS:
        fk = em.forces(far=False, intra=False, wall=False, graphite=False, kernels=[k])
        names.append(k)
        rows = np.nonzero(np.abs(fk).max(axis=1))[0].astype(np.int32)
        out['Ki_' + k] = rows
        out['Kv_' + k] = fk[rows]
    out['kernels'] = np.array(names)
    # one full step: init_accel, part A, rebuild, forces, part B
    em.init_accel(out['F'])
    out['acc0'] = em.acc.copy()
    sums = em.step_a()
    out['A_pos'], out['A_vel'], out['A_eta'], out['A_sums'] = em.pos.copy(), em.vel.copy(), \
        em.eta, sums
    em.reclassify()
    F1 = em.forces()
    out['F1'] = F1
    em.step_b(F1)
    out['B_vel'], out['B_acc'] = em.vel.copy(), em.acc.copy()
    np.savez_compressed(os.path.join(HERE, 'small_system.npz'), **out)


def medium(ref):
    s = medium_system()
    em = rh.Emulator(ref, s)
    np.savez_compressed(os.path.join(HERE, 'medium_total.npz'), n_water=s.n_water,
                        n_surf=s.n_surf, pos=s.pos, vel=s.vel, graphite=s.graphite,
                        F=em.forces())


R2_EVERY = 7


def r2(ref):
    from surfactant_enhanced_spreading_amd import builder
    s0 = builder.r2_system()
    with tempfile.TemporaryDirectory() as d:
        path = os.path.join(d, 'data_cont2.dat')
        fmt.write_config(path, s0)
        s = fmt.read_config(path, s0.n_water, s0.n_surf)
        # the file's own fields, as integers in units of its last digit
        vals = np.array([[fmt.parse_f(f, 3) for f in fmt._split(l.rstrip('\n'), (8, 9, 9))]
                         for l in open(path)])
    s.graphite = s0.graphite
    pos_milli = np.rint(vals[0::2] * 1000.0).astype(np.int32)
    vel_milli = np.rint(vals[1::2] * 1000.0).astype(np.int32)
    em = rh.Emulator(ref, s)
    F = em.forces()
    em.init_accel(F)
    em.step_a()
    em.reclassify()
    F1 = em.forces()
    em.step_b(F1)
    k = np.arange(0, s.n_sites, R2_EVERY)
    np.savez_compressed(
        os.path.join(HERE, 'r2_system.npz'), n_water=s.n_water, n_surf=s.n_surf,
        pos_milli=pos_milli, vel_milli=vel_milli,
        graphite_make=True, every=R2_EVERY, sites=k, F=F[k], A_pos=em.pos[k], A_eta=em.eta,
        F1=F1[k], B_vel=em.vel[k], B_acc=em.acc[k])


def sha(path):
    with open(path, 'rb') as fh:
        return hashlib.sha256(fh.read()).hexdigest()


def stored_digest(ref, em):
    """SHA-256 over the rows (x, y, z, vx, vy, vz) that read_config stores in
    the own boxes of all 27 ranks, rows sorted, as little-endian float64."""
    rows = []
    cwd = os.getcwd()
    for rk_i, rk in enumerate(em.ranks()):
        em.load_rank(rk)
        for c, tag in enumerate(('w1', 'w2', 's')):
            ref['ictw1' if c == 0 else 'ictw2' if c == 1 else 'icts'][:] = 0
            for a in 'xyz':
                ref['r%s%s' % (a, tag)][:] = 0.0
                ref['v%s%s' % (a, tag)][:] = 0.0
        os.chdir(em.workdir)
        try:
            ref.call('read_config', rk_i)
        finally:
            os.chdir(cwd)
        for c, tag in enumerate(('w1', 'w2', 's')):
            grid, halo = em.RING[c]
            nsite = 29 if c == 2 else 3
            cnt = ref['ictw1' if c == 0 else 'ictw2' if c == 1 else 'icts']
            for lz in range(halo, halo + 4):
                for ly in range(halo, halo + 4):
                    for lx in range(halo, halo + 4):
                        box = lx + grid * ly + grid * grid * lz
                        b64 = (lx - halo) + 4 * (ly - halo) + 16 * (lz - halo)
                        k = cnt[box] * nsite
                        if k == 0:
                            continue
                        r = [ref['r%s%s' % (a, tag)][:k, box] for a in 'xyz']
                        v = [ref['v%s%s' % (a, tag)][:k, b64] for a in 'xyz']
                        rows.append(np.stack(r + v, axis=1))
    a = np.concatenate(rows)
    a = a[np.lexsort(a.T[::-1])]
    return hashlib.sha256(np.ascontiguousarray(a, dtype='<f8').tobytes()).hexdigest(), len(a)


def formats(ref):
    out = {}
    with tempfile.TemporaryDirectory() as d:
        subprocess.run([rh.MAKE_GRAPHITE], cwd=d, check=True, capture_output=True)
        g = os.path.join(d, 'graphite.dat')
        lines = open(g).read().split('\n')
        out['graphite'] = dict(sha256=sha(g), lines=len(lines) - 1, first=lines[0], last=lines[-2])
        s = file_format_system()
        ref.write_pdb(s, 42, d)
        _, sname = fmt.pdb_names(42, s.n_water)
        out['pdb'] = {f: sha(os.path.join(d, f)) for f in sorted(os.listdir(d)) if f.endswith('.pdb')}
        fmt.write_config(os.path.join(d, 'data_cont2.dat'), s)
        out['data_cont2_sha256'] = sha(os.path.join(d, 'data_cont2.dat'))
        em = rh.Emulator(ref, fmt.read_config(os.path.join(d, 'data_cont2.dat'), s.n_water, s.n_surf))
        em.workdir = d
        out['read_config_sha256'], out['read_config_rows'] = stored_digest(ref, em)
    with open(os.path.join(HERE, 'formats.json'), 'w') as fh:
        json.dump(out, fh, indent=1, sort_keys=True)
        fh.write('\n')


def velocities(ref):
    import ctypes
    f = ref.lib.random_
    f.argtypes = [ctypes.POINTER(ctypes.c_double)] * 2
    x, seq = 0.378389247, []
    for _ in range(64):
        a, b = ctypes.c_double(x), ctypes.c_double()
        f(ctypes.byref(a), ctypes.byref(b))
        x = b.value
        seq.append(x.hex())
    with open(os.path.join(HERE, 'velocities.json'), 'w') as fh:
        json.dump(dict(seed=0.378389247, lcg=seq), fh, indent=1)
        fh.write('\n')


if __name__ == '__main__':
    ref = rh.RefLib()
    if len(sys.argv) > 1 and sys.argv[1] in ('formats', 'velocities', 'r2'):
        {'formats': formats, 'velocities': velocities, 'r2': r2}[sys.argv[1]](ref)
        sys.exit(0)
    formats(ref)
    velocities(ref)
    tables(ref)
    ylm(ref)
    small(ref)
    medium(ref)
    r2(ref)
    for f in sorted(os.listdir(HERE)):
        if f.endswith('.npz'):
            print(f, os.path.getsize(os.path.join(HERE, f)))
