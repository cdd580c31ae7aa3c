"""Start-up velocities and temperature diagnostics (host side).

* ``reference_lcg`` / ``init_temp`` -- the reference's own initialiser replayed
  exactly (``src/init_temp.f90:21-822`` with the generator of
  ``src/random.f90:1-9``): every rank restarts the generator from the seed
  0.378389247 and walks its own L4 boxes (W1, then W2, then surfactant
  molecules of each box, in storage order), giving every site a random *unit*
  vector (not a Maxwell-Boltzmann draw, despite the reference's message);
  each rank removes its own mean per site group (O, H, CH2, Si, CH3); one
  global factor then sets the temperature to ``tr``.  The velocities are not
  projected onto the constraints, as in the reference.
* ``maxwell_boltzmann`` + ``project_constraints`` -- the physical alternative:
  seeded Gaussian velocities, projected onto the bond constraints (the
  velocity half of RATTLE, iterated to convergence), rescaled to ``tr``.
* ``temperature`` (``src/compute_temperature.f90:240-244``) and
  ``molecule_counts`` (``src/get_molecule_count.f90``).

Velocities use the reference's storage convention, v*dt.  Classes and
storage boxes come from the engine (``Engine.classes()``): class 0 W1, 1 W2,
2 surfactant; ``anc`` the global L4 box index x + nb (y + nb z).
"""
import numpy as np

from . import forcefield as ff

LCG_SEED = 0.378389247                     # src/init_temp.f90:21
_LCG_K, _LCG_J, _LCG_M = 5701, 3612, 566927  # src/random.f90:4
# site group per surfactant site: 0 O, 1 H, 2 CH2, 3 Si, 4 CH3 (the sums of
# src/init_temp.f90:146-400); waters are O, H, H
_SURF_GROUP = np.array([1, 0, 2, 2, 0, 2, 2, 0, 2, 2, 0, 2, 2, 0, 2, 2, 2, 3, 4, 0,
                        3, 4, 4, 4, 0, 3, 4, 4, 4])
_WATER_GROUP = np.array([0, 1, 1])


def reference_lcg(x):
    """One step of the reference's generator: x in (0, 1) -> next value."""
    ix = int(x * float(_LCG_M))
    irand = (_LCG_J * ix + _LCG_K) % _LCG_M
    return (float(irand) + 0.5) / float(_LCG_M)


def _df(n_water, n_surf):
    return 3 * (3 * n_water + 29 * n_surf) - (3 * n_water + 28 * n_surf)


def site_groups(n_water, n_surf):
    """Group of every site (0 O, 1 H, 2 CH2, 3 Si, 4 CH3)."""
    return np.concatenate([np.tile(_WATER_GROUP, n_water), np.tile(_SURF_GROUP, n_surf)])


def group_masses():
    """The mass each group is weighted with (src/init_temp.f90:560: emw(1),
    emw(2), ems(3), ems(18), ems(19))."""
    return np.array([ff.EMW[0], ff.EMW[1], ff.EMS[2], ff.EMS[17], ff.EMS[18]])


def _storage_order(n_water, n_surf, cls, anc, nb):
    """Per rank, the molecules in the order init_temp visits them: own boxes
    ib4 = 1..64 (x fastest), in each W1, then W2, then surfactants, each in
    global order (the order read_config files them in)."""
    nm = n_water + n_surf
    ax, ay, az = anc % nb, (anc // nb) % nb, anc // (nb * nb)
    rank = (ax // 4) + (nb // 4) * ((ay // 4) + (nb // 4) * (az // 4))
    ib4 = (ax % 4) + 4 * (ay % 4) + 16 * (az % 4)
    key = np.lexsort((np.arange(nm), cls, ib4))  # by box, then class, then molecule
    return [key[rank[key] == r] for r in range((nb // 4) ** 3)]


def init_temp(sys_, cls, anc, tr=ff.TR_DEFAULT, nb=12):
    """The reference's start-up velocities for sys_ (filed by cls, anc):
    returns (N, 3) v*dt."""
    nw, ns = sys_.n_water, sys_.n_surf
    first = np.concatenate([3 * np.arange(nw), 3 * nw + 29 * np.arange(ns)]).astype(np.int64)
    size = np.concatenate([np.full(nw, 3), np.full(ns, 29)])
    grp = site_groups(nw, ns)
    v = np.zeros((sys_.n_sites, 3))
    for mols in _storage_order(nw, ns, np.asarray(cls), np.asarray(anc), nb):
        if len(mols) == 0:
            continue
        x = LCG_SEED
        sites = np.concatenate([np.arange(first[m], first[m] + size[m]) for m in mols])
        for i in sites:
            d = []
            for _ in range(3):
                x = reference_lcg(x)
                d.append(2.0 * (x - 0.5))
            inv = 1.0 / np.sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2])
            v[i] = (d[0] * inv, d[1] * inv, d[2] * inv)
        # this rank's mean per group (src/init_temp.f90:122-400), then removed
        for g in range(5):
            sel = sites[grp[sites] == g]
            if len(sel):
                v[sel] -= v[sel].sum(axis=0) * (1.0 / len(sel))
    mv2 = (group_masses()[grp] * (v * v).sum(axis=1)).sum()
    temp = mv2 / ff.DT ** 2 / _df(nw, ns)
    return v * np.sqrt(tr / temp)


def temperature(sys_, vel=None):
    """sum m v^2 / dt^2 / df with the group masses (src/compute_temperature.f90:240-244)."""
    v = sys_.vel if vel is None else vel
    grp = site_groups(sys_.n_water, sys_.n_surf)
    mv2 = (group_masses()[grp] * (v * v).sum(axis=1)).sum()
    return mv2 / ff.DT ** 2 / _df(sys_.n_water, sys_.n_surf)


def molecule_counts(cls):
    """(W1, W2, surfactant) molecule counts and the site total
    (src/get_molecule_count.f90)."""
    cls = np.asarray(cls)
    w1, w2, s = int((cls == 0).sum()), int((cls == 1).sum()), int((cls == 2).sum())
    return dict(w1=w1, w2=w2, surfactant=s, sites=3 * (w1 + w2) + 29 * s)


def _bonds(sys_):
    """All constraint pairs as (site a, site b, d^2) arrays: the water triangle
    (src/init_arrays.f90:566-574) and the 28 surfactant bonds."""
    nw, ns = sys_.n_water, sys_.n_surf
    out_a, out_b, out_d = [], [], []
    for c in range(3):
        ia, ib = c, (c + 1) % 3
        out_a.append(3 * np.arange(nw) + ia)
        out_b.append(3 * np.arange(nw) + ib)
        out_d.append(np.full(nw, ff.DSQW[c]))
    base = 3 * nw + 29 * np.arange(ns)
    for c in range(28):
        out_a.append(base + ff.IBOND[c] - 1)
        out_b.append(base + ff.JBOND[c] - 1)
        out_d.append(np.full(ns, ff.DSQS[c]))
    return out_a, out_b, out_d


def project_constraints(sys_, vel, tol=1e-13, max_sweeps=1000):
    """Remove the velocity components along every constraint (r_ab . v_ab = 0)
    by sweeping the bonds in order, as the reference's RATTLE does, until the
    largest relative violation is below tol."""
    v = np.array(vel, dtype=np.float64, copy=True)
    rm = 1.0 / sys_.masses()
    ia, ib, dsq = _bonds(sys_)
    for _ in range(max_sweeps):
        worst = 0.0
        for a, b, d in zip(ia, ib, dsq):
            if len(a) == 0:
                continue
            rab = sys_.pos[a] - sys_.pos[b]
            vab = v[a] - v[b]
            rv = (rab * vab).sum(axis=1)
            g = -rv / ((rm[a] + rm[b]) * d)
            v[a] += (rm[a] * g)[:, None] * rab
            v[b] -= (rm[b] * g)[:, None] * rab
            scale = np.sqrt(d) * np.maximum(np.abs(vab).max(axis=1), 1e-300)
            worst = max(worst, float((np.abs(rv) / scale).max()))
        if worst < tol:
            return v
    raise RuntimeError('constraint projection did not converge')


def maxwell_boltzmann(sys_, tr=ff.TR_DEFAULT, seed=0):
    """Gaussian velocities at temperature tr, projected onto the constraints
    and rescaled to tr (v*dt units)."""
    rng = np.random.default_rng(seed)
    m = sys_.masses()
    v = rng.normal(size=(sys_.n_sites, 3)) * (np.sqrt(tr / m) * ff.DT)[:, None]
    v = project_constraints(sys_, v)
    return v * np.sqrt(tr / temperature(sys_, v))
