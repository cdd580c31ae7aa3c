"""Synthetic system builder: SPC/E water droplets carrying 29-site M(D'E4OH)M
surfactants above a two-layer graphite sheet, in the reference's reduced units.

Site layout (the order the reference reads its configuration in,
``src/read_config.f90:19-34``): all waters first, each as (O, H1, H2), then
all surfactants, 29 consecutive sites each (site 14 is the oxygen the
reference files the molecule under, ``src/read_config.f90:145-175``).

Velocities use the reference's storage convention: v*dt (displacement per
step); accelerations are stored as F*dt^2/(2m) (``src/init_accel.f90:21``).
"""
import math
from dataclasses import dataclass, field

import numpy as np

from . import forcefield as ff

WATER_DENSITY_A3 = 0.0334  # molecules per cubic Angstrom


@dataclass
class System:
    n_water: int
    n_surf: int
    pos: np.ndarray            # (N, 3) float64, sigma units
    vel: np.ndarray            # (N, 3) v*dt
    acc: np.ndarray = None     # (N, 3) F*dt^2/(2m)
    graphite: np.ndarray = field(default_factory=lambda: np.zeros((0, 3)))

    @property
    def n_sites(self):
        return 3 * self.n_water + 29 * self.n_surf

    def types(self):
        """Site type in the 32-type space (0..2 water, 3..31 surfactant)."""
        t = np.empty(self.n_sites, dtype=np.int32)
        t[:3 * self.n_water] = np.tile(np.arange(3, dtype=np.int32), self.n_water)
        t[3 * self.n_water:] = np.tile(np.arange(3, 32, dtype=np.int32), self.n_surf)
        return t

    def masses(self):
        return ff.type_tables()['mass'][self.types()]

    def copy(self):
        return System(self.n_water, self.n_surf, self.pos.copy(), self.vel.copy(),
                      None if self.acc is None else self.acc.copy(),
                      self.graphite.copy())


def _rot_random(rng):
    """Uniform random rotation matrix (from a random unit quaternion)."""
    q = rng.normal(size=4)
    q /= np.linalg.norm(q)
    a, b, c, d = q
    return np.array([
        [a * a + b * b - c * c - d * d, 2 * (b * c - a * d), 2 * (b * d + a * c)],
        [2 * (b * c + a * d), a * a - b * b + c * c - d * d, 2 * (c * d - a * b)],
        [2 * (b * d - a * c), 2 * (c * d + a * b), a * a - b * b - c * c + d * d]])


def water_template():
    """Rigid water with |OH| = 1 A and HOH = 109.47 deg, the constraint
    geometry of ``src/init_arrays.f90:566-574``."""
    half = 0.5 * 109.47 * math.pi / 180.0
    r = 1.0 / ff.SIGMA
    return np.array([[0.0, 0.0, 0.0],
                     [r * math.sin(half), r * math.cos(half), 0.0],
                     [-r * math.sin(half), r * math.cos(half), 0.0]])


def _place(a, b, c, bond, angle, dih):
    """Position of a new site bonded to a, with angle (new, a, b) and dihedral
    (new, a, b, c) -- the usual z-matrix construction."""
    bc = b - c
    ab = a - b
    n = np.cross(bc, ab)
    if np.linalg.norm(n) < 1e-12:
        n = np.cross(ab, [1.0, 0.3, 0.1])
    n /= np.linalg.norm(n)
    u = ab / np.linalg.norm(ab)
    m = np.cross(n, u)
    d2 = np.array([-bond * math.cos(angle),
                   bond * math.sin(angle) * math.cos(dih),
                   bond * math.sin(angle) * math.sin(dih)])
    return a + d2[0] * u + d2[1] * m + d2[2] * n


def _bond_graph_distance():
    adj = {i: set() for i in range(1, 30)}
    for b in range(28):
        adj[ff.IBOND[b]].add(ff.JBOND[b])
        adj[ff.JBOND[b]].add(ff.IBOND[b])
    dist = np.full((30, 30), 99, dtype=int)
    for s in range(1, 30):
        dist[s, s] = 0
        front = [s]
        while front:
            nxt = []
            for a in front:
                for c in adj[a]:
                    if dist[s, c] > dist[s, a] + 1:
                        dist[s, c] = dist[s, a] + 1
                        nxt.append(c)
            front = nxt
    return dist[1:, 1:]


def surfactant_template():
    """One extended M(D'E4OH)M conformer whose 28 bond lengths equal the
    constraint lengths exactly and whose angles sit at their equilibria; the
    siloxane-head dihedrals are picked on a small grid to maximise the closest
    approach of sites three or more bonds apart.  (Computed once; a copy is
    returned.)"""
    global _TEMPLATE
    if _TEMPLATE is None:
        _TEMPLATE = _surfactant_template()
    return _TEMPLATE.copy()


_TEMPLATE = None


def _surfactant_template():
    blen = {}
    for b in range(28):
        i, j = ff.IBOND[b], ff.JBOND[b]
        blen[(i, j)] = blen[(j, i)] = math.sqrt(ff.DSQS[b])
    ang = {}
    for a in range(36):
        i, j, k = ff.ANG_I[a], ff.ANG_J[a], ff.ANG_K[a]
        ang[(i, j, k)] = ang[(k, j, i)] = ff.TH[a]
    tet = 109.5 * math.pi / 180.0

    def A(i, j, k):
        return ang.get((i, j, k), tet)

    x = {}
    x[1] = np.zeros(3)
    x[2] = np.array([blen[(1, 2)], 0.0, 0.0])
    th = A(1, 2, 3)
    x[3] = x[2] + blen[(2, 3)] * np.array([-math.cos(th), math.sin(th), 0.0])
    for s in range(4, 19):  # extended chain, all trans
        x[s] = _place(x[s - 1], x[s - 2], x[s - 3], blen[(s - 1, s)],
                      A(s, s - 1, s - 2), math.pi)
    far = _bond_graph_distance() >= 3
    third = 2 * math.pi / 3

    def head(d0, d1, d2):
        y = dict(x)
        y[19] = _place(y[18], y[17], y[16], blen[(18, 19)], A(19, 18, 17), d0)
        y[20] = _place(y[18], y[17], y[16], blen[(18, 20)], A(20, 18, 17), d0 + third)
        y[25] = _place(y[18], y[17], y[16], blen[(18, 25)], A(25, 18, 17), d0 - third)
        y[21] = _place(y[20], y[18], y[17], blen[(20, 21)], A(21, 20, 18), d1)
        for s, d in ((22, 0.0), (23, third), (24, -third)):
            y[s] = _place(y[21], y[20], y[18], blen[(21, s)], A(s, 21, 20), math.pi + d)
        y[26] = _place(y[25], y[18], y[17], blen[(25, 26)], A(26, 25, 18), d2)
        for s, d in ((27, 0.0), (28, third), (29, -third)):
            y[s] = _place(y[26], y[25], y[18], blen[(26, s)], A(s, 26, 25), math.pi + d)
        X = np.array([y[s] for s in range(1, 30)])
        D = np.sqrt(((X[:, None] - X[None]) ** 2).sum(-1))
        return D[far].min(), X

    best = None
    grid = np.linspace(-math.pi, math.pi, 12, endpoint=False)
    for d0 in grid:
        for d1 in grid:
            for d2 in grid:
                dmin, X = head(d0, d1, d2)
                if best is None or dmin > best[0]:
                    best = (dmin, X)
    return best[1]


def _align(v_from, v_to):
    """Rotation matrix taking unit vector v_from onto unit vector v_to."""
    a = v_from / np.linalg.norm(v_from)
    b = v_to / np.linalg.norm(v_to)
    v = np.cross(a, b)
    c = float(np.dot(a, b))
    if c < -0.999999:
        p = np.cross(a, [1.0, 0.0, 0.0])
        if np.linalg.norm(p) < 1e-6:
            p = np.cross(a, [0.0, 1.0, 0.0])
        p /= np.linalg.norm(p)
        return 2.0 * np.outer(p, p) - np.eye(3)
    vx = np.array([[0, -v[2], v[1]], [v[2], 0, -v[0]], [-v[1], v[0], 0]])
    return np.eye(3) + vx + vx @ vx / (1.0 + c)


def _surfactants_rsa(n_surf, r_drop, c, body0, axis, length, rng, min_ss, hemisphere=False):
    """n_surf surfactants on the droplet's surface, placed one by one: evenly
    spread anchor directions (a Fibonacci sphere, in random order), and per
    chain 16 spins about its axis x 8 anchor radii (from the radius droplet()
    uses, 0.25 sigma apart, the chain pointing outward), the first of them in
    that order whose sites all lie at least min_ss from every site placed
    before (the last radius if none)."""
    from scipy.spatial import cKDTree
    out = np.zeros((n_surf, 29, 3))
    r0 = max(r_drop + 0.35 * length - length, 0.75 * r_drop)
    order = rng.permutation(n_surf)
    spins = rng.uniform(0, 2 * math.pi, size=(n_surf, 16))
    tree, ntree = None, 0  # tree over the first ntree chains; later ones checked directly
    for k, s in enumerate(order):
        zz = 1.0 - 2.0 * (s + 0.5) / n_surf
        if hemisphere:
            zz = 1.0 - (s + 0.5) / n_surf
        phi = s * math.pi * (3.0 - math.sqrt(5.0))
        d = np.array([math.sqrt(max(0.0, 1 - zz * zz)) * math.cos(phi),
                      math.sqrt(max(0.0, 1 - zz * zz)) * math.sin(phi), zz])
        R = _align(axis, d)
        K = np.array([[0, -d[2], d[1]], [d[2], 0, -d[0]], [-d[1], d[0], 0]])
        base = (R @ body0.T).T
        cands = []
        for step in range(8):
            for ang in spins[s]:
                spin = np.eye(3) + math.sin(ang) * K + (1 - math.cos(ang)) * K @ K
                cands.append((spin @ base.T).T + c + d * (r0 + 0.25 * step))
        cands = np.array(cands)  # (128, 29, 3)
        dmin = np.full(len(cands), np.inf)
        if tree is not None:
            dd, _ = tree.query(cands.reshape(-1, 3), k=1)
            dmin = dd.reshape(len(cands), 29).min(axis=1)
        if k > ntree:
            recent = out[order[ntree:k]].reshape(-1, 3)
            dr = np.sqrt(((cands.reshape(-1, 1, 3) - recent[None]) ** 2).sum(axis=2)).min(axis=1)
            dmin = np.minimum(dmin, dr.reshape(len(cands), 29).min(axis=1))
        ok = np.nonzero(dmin >= min_ss)[0]
        out[s] = cands[ok[0]] if len(ok) else cands[-1]
        if k + 1 - ntree >= 8:
            ntree = k + 1
            tree = cKDTree(out[order[:ntree]].reshape(-1, 3))
    return out


def droplet(n_water, n_surf, center_A=(0.0, 0.0, 0.0), seed=0, min_sep_A=2.6, hemisphere=False,
            spread='fibonacci', min_ss_A=3.0):
    """Positions of one droplet: waters on a simple-cubic lattice at liquid
    density with random orientations, cut to the n_water lattice sites closest
    to the centre, and n_surf surfactants spread over the droplet surface with
    the EO chain (site 1 end) pointing inward and the trisiloxane outward.

    hemisphere: only lattice sites at or above the centre's height (a
    hemispherical droplet whose flat face lies at center_A's z), radius from
    the hemisphere's volume; surfactants on the curved surface only.

    spread: 'fibonacci' anchors the surfactants at evenly spread directions
    with one random spin each (fine while they are few); 'rsa' places them
    one by one, choosing spin and anchor radius so that each lies at least
    min_ss_A from the ones before (dense coverage).

    Returns (water_pos (n_water,3,3), surf_pos (n_surf,29,3)) in sigma units.
    """
    rng = np.random.default_rng(seed)
    a = (1.0 / WATER_DENSITY_A3) ** (1.0 / 3.0) / ff.SIGMA
    c = np.asarray(center_A, dtype=float) / ff.SIGMA
    vol = n_water / WATER_DENSITY_A3 * (2.0 if hemisphere else 1.0)
    r_drop = (3.0 * vol / (4.0 * math.pi)) ** (1.0 / 3.0) / ff.SIGMA
    tmpl = surfactant_template()
    axis = tmpl[17] - tmpl[0]
    length = np.linalg.norm(axis)
    surf = np.zeros((n_surf, 29, 3))
    if spread == 'rsa':
        surf = _surfactants_rsa(n_surf, r_drop, c, tmpl - tmpl[0], axis, length, rng,
                                min_ss_A / ff.SIGMA, hemisphere)
    for s in range(n_surf if spread != 'rsa' else 0):
        # Fibonacci sphere for evenly spread anchor directions
        zz = 1.0 - 2.0 * (s + 0.5) / max(n_surf, 1)
        if hemisphere:
            zz = 1.0 - (s + 0.5) / max(n_surf, 1)  # upper half only
        phi = s * math.pi * (3.0 - math.sqrt(5.0))
        d = np.array([math.sqrt(max(0.0, 1 - zz * zz)) * math.cos(phi),
                      math.sqrt(max(0.0, 1 - zz * zz)) * math.sin(phi), zz])
        R = _align(axis, d)
        spin = _align(d, d)
        ang = rng.uniform(0, 2 * math.pi)
        K = np.array([[0, -d[2], d[1]], [d[2], 0, -d[0]], [-d[1], d[0], 0]])
        spin = np.eye(3) + math.sin(ang) * K + (1 - math.cos(ang)) * K @ K
        body = (spin @ (R @ (tmpl - tmpl[0]).T)).T
        # site 1 a little inside the surface, siloxane head outside
        start = c + d * max(r_drop + 0.35 * length - length, 0.75 * r_drop)
        surf[s] = body + start
    # lattice sites inside a generous sphere, closest to the centre first
    nmax = int(math.ceil((r_drop * 1.6) / a)) + 2
    g = np.arange(-nmax, nmax + 1) * a
    X, Y, Z = np.meshgrid(g, g, g, indexing='ij')
    pts = np.stack([X.ravel(), Y.ravel(), Z.ravel()], axis=1)
    pts += rng.uniform(-0.02, 0.02, size=3)
    if hemisphere:
        pts = pts[pts[:, 2] >= 0.0]
    d2 = np.sum(pts * pts, axis=1)
    pts = pts[np.argsort(d2, kind='stable')] + c
    # candidates: enough for the waters plus the lattice points the
    # surfactant sites exclude (the first n_water survivors are the same for
    # any larger pool)
    pts = pts[:int(n_water * 1.5) + 64 + 60 * n_surf]
    if n_surf:
        from scipy.spatial import cKDTree
        near = cKDTree(surf.reshape(-1, 3)).query(pts, k=1)[0]
        pts = pts[near > min_sep_A / ff.SIGMA]
    if len(pts) < n_water:
        raise ValueError("lattice too small for the requested water count")
    pts = pts[:n_water]
    w = water_template()
    q = rng.normal(size=(n_water, 4))
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    qa, qb, qc, qd = q.T
    R = np.stack([
        np.stack([qa * qa + qb * qb - qc * qc - qd * qd, 2 * (qb * qc - qa * qd),
                  2 * (qb * qd + qa * qc)], axis=1),
        np.stack([2 * (qb * qc + qa * qd), qa * qa - qb * qb + qc * qc - qd * qd,
                  2 * (qc * qd - qa * qb)], axis=1),
        np.stack([2 * (qb * qd - qa * qc), 2 * (qc * qd + qa * qb),
                  qa * qa - qb * qb - qc * qc + qd * qd], axis=1)], axis=1)
    waters = np.einsum('kij,sj->ksi', R, w) + pts[:, None, :]
    return waters, surf


def graphite_sheet(half_side_A=60.0, z_top=-16.0, layer_gap_A=3.41, ccbond_A=1.421, layers=2):
    """AB-stacked graphene layers of C atoms (sigma units), the top layer at
    z = z_top sigma and each further layer layer_gap_A below, like the sheet
    the reference's own generator writes (``src/make_graphite.f90:14-80``,
    two layers).  Coordinates are rounded to 1e-4 sigma, the precision of the
    reference's graphite file (``src/read_graphite.f90:216``)."""
    a = ccbond_A / ff.SIGMA
    h = half_side_A / ff.SIGMA
    ax = np.array([1.5 * a, math.sqrt(3) / 2 * a])
    ay = np.array([0.0, math.sqrt(3) * a])
    n = int(h / a) + 3
    k = np.arange(-n, n + 1, dtype=float)
    out = []
    for layer in range(layers):
        zl = z_top - layer * layer_gap_A / ff.SIGMA
        shift = np.array([a, 0.0]) * (layer % 2)
        # lattice points in (i, j, basis atom) order
        base = k[:, None, None] * ax + k[None, :, None] * ay + shift      # (i, j, 2)
        pts = (base[:, :, None, :] + np.array([[0.0, 0.0], [a, 0.0]])[None, None]).reshape(-1, 2)
        pts = pts[(np.abs(pts[:, 0]) <= h) & (np.abs(pts[:, 1]) <= h)]
        out.append(np.column_stack([pts, np.full(len(pts), zl)]))
    g = np.round(np.concatenate(out, axis=0), 4)
    return g


def maxwell_velocities(masses, tr, seed):
    """Gaussian velocities at reduced temperature tr in the v*dt convention
    (per degree of freedom m <v^2> = tr dt^2, cf. ``src/compute_temperature.f90:242``)."""
    rng = np.random.default_rng(seed)
    sd = np.sqrt(tr / masses)[:, None] * ff.DT
    return rng.normal(size=(len(masses), 3)) * sd


def assemble(droplets, graphite=None, tr=ff.TR_DEFAULT, seed=0):
    """Concatenate droplets (lists of (waters, surfs)) into one System with all
    waters first, then all surfactants."""
    ws = [d[0] for d in droplets if len(d[0])]
    ss = [d[1] for d in droplets if len(d[1])]
    W = np.concatenate(ws, axis=0) if ws else np.zeros((0, 3, 3))
    S = np.concatenate(ss, axis=0) if ss else np.zeros((0, 29, 3))
    pos = np.concatenate([W.reshape(-1, 3), S.reshape(-1, 3)], axis=0)
    sys_ = System(len(W), len(S), pos, np.zeros_like(pos),
                  graphite=np.zeros((0, 3)) if graphite is None else graphite)
    sys_.vel = maxwell_velocities(sys_.masses(), tr, seed + 7)
    return sys_


def bench_system(n_water=20000, n_surf=200, seed=20261204):
    """The throughput workload: one droplet of n_water SPC/E waters carrying
    n_surf surfactants, centred in the reference's 144 A cube above the sheet."""
    d = droplet(n_water, n_surf, center_A=(0.0, 0.0, 4.0), seed=seed)
    return assemble([d], graphite=graphite_sheet(), seed=seed)


def config1_system(n_water=20000, n_surf=200, seed=20261204):
    """The 20k-water + 200-surfactant droplet in the 16^3-box domain (nb2 = 4:
    4,096 L4 boxes of 12 A, a 192 A cube): the droplet of bench_system above a
    sheet spanning the domain's 192 x 192 A floor."""
    d = droplet(n_water, n_surf, center_A=(0.0, 0.0, 4.0), seed=seed)
    return assemble([d], graphite=graphite_sheet(half_side_A=96.0), seed=seed)


def c0_system(n_water=2000, seed=20261203):
    """2,000 waters as a hemispherical droplet (radius ~31 A) resting 3.2 A
    above a three-layer graphite slab, in the 8^3-box domain (nb2 = 2, a 96 A
    cube): the slab's top layer at z = -41 A, inside the bottom layer of L4
    boxes, so the droplet's lowest sites feel the graphite."""
    z_top = -41.0
    d = droplet(n_water, 0, center_A=(0.0, 0.0, z_top + 3.2), seed=seed, hemisphere=True)
    g = graphite_sheet(half_side_A=46.0, z_top=z_top / ff.SIGMA, layer_gap_A=3.35, layers=3)
    return assemble([d], graphite=g, seed=seed)


def film(n_water, n_surf, half_side_A, z_base_A, seed=0, min_sep_A=2.6):
    """A water film on a simple-cubic lattice at liquid density over
    |x|, |y| <= half_side_A from z_base_A up, with n_surf upright surfactants
    (site 1 down, the trisiloxane head up, all turned alike): half of them on a
    square grid whose heads stand above the film's top face, the rest on the
    same grid shifted by half a cell, within the film at mid-depth.  The grid
    is wide enough that no two heads meet; waters within min_sep_A of a
    surfactant site are left out.  Returns (water_pos, surf_pos) in sigma
    units."""
    rng = np.random.default_rng(seed)
    a_A = (1.0 / WATER_DENSITY_A3) ** (1.0 / 3.0)
    nx = int(2 * half_side_A / a_A)
    tmpl = surfactant_template()
    axis = tmpl[17] - tmpl[0]
    length = np.linalg.norm(axis)
    R = _align(axis, np.array([0.0, 0.0, 1.0]))
    body0 = (R @ (tmpl - tmpl[0]).T).T
    n_top = (n_surf + 1) // 2
    m = max(1, int(math.ceil(math.sqrt(max(n_top, 1)))))
    step = 2 * half_side_A / m / ff.SIGMA
    # film thick enough for the waters plus the volume the surfactants take
    vol_s = 29 * 30.0  # A^3 per surfactant, roughly
    thick = (n_water / WATER_DENSITY_A3 + n_surf * vol_s) / (4 * half_side_A ** 2)
    nz = int(math.ceil(thick / a_A)) + 2
    gx = (np.arange(nx) - (nx - 1) / 2.0) * a_A
    gz = z_base_A + (np.arange(nz) + 0.5) * a_A
    X, Y, Z = np.meshgrid(gx, gx, gz, indexing='ij')
    pts = np.stack([X.ravel(), Y.ravel(), Z.ravel()], axis=1) / ff.SIGMA
    pts += rng.uniform(-0.02, 0.02, size=3)
    top = (z_base_A + thick) / ff.SIGMA
    mid = (z_base_A + 0.5 * thick) / ff.SIGMA - 0.5 * length
    surf = np.zeros((n_surf, 29, 3))
    for s_ in range(n_surf):
        lay, k = (0, s_) if s_ < n_top else (1, s_ - n_top)
        off = 0.5 * lay
        x0 = -half_side_A / ff.SIGMA + (k % m + 0.5 + off) * step
        y0 = -half_side_A / ff.SIGMA + (k // m % m + 0.5 + off) * step
        z0 = top - 0.3 * length if lay == 0 else mid
        surf[s_] = body0 + np.array([x0, y0, z0])
    if n_surf:
        from scipy.spatial import cKDTree
        near = cKDTree(surf.reshape(-1, 3)).query(pts, k=1)[0]
        pts = pts[near > min_sep_A / ff.SIGMA]
    pts = pts[np.argsort(pts[:, 2], kind='stable')]
    if len(pts) < n_water:
        raise ValueError("film lattice too small for the requested water count")
    pts = pts[:n_water]
    w = water_template()
    q = rng.normal(size=(n_water, 4))
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    qa, qb, qc, qd = q.T
    Rq = np.stack([
        np.stack([qa * qa + qb * qb - qc * qc - qd * qd, 2 * (qb * qc - qa * qd),
                  2 * (qb * qd + qa * qc)], axis=1),
        np.stack([2 * (qb * qc + qa * qd), qa * qa - qb * qb + qc * qc - qd * qd,
                  2 * (qc * qd - qa * qb)], axis=1),
        np.stack([2 * (qb * qd - qa * qc), 2 * (qc * qd + qa * qb),
                  qa * qa - qb * qb - qc * qc + qd * qd], axis=1)], axis=1)
    waters = np.einsum('kij,sj->ksi', Rq, w) + pts[:, None, :]
    return waters, surf


def c4_system(n_water=200000, n_surf=4000, seed=20261207):
    """200,000 waters with 4,000 surfactants as a film 400 x 400 A wide on a
    two-layer graphite slab, in the 36^3-box domain (nb2 = 9, a 432 A cube:
    the smallest cube of 12 A boxes that holds the film's width); the slab's
    top layer at z = -209 A, inside the bottom layer of L4 boxes."""
    z_top = -209.0
    d = film(n_water, n_surf, 200.0, z_top + 3.2, seed=seed)
    g = graphite_sheet(half_side_A=206.0, z_top=z_top / ff.SIGMA, layer_gap_A=3.41)
    return assemble([d], graphite=g, seed=seed)


def ensemble_systems(n=64, seed=20261208, n_min=500, n_max=30000, surf_wt_max=0.05, layers=2):
    """The batched-ensemble workload: n independent droplets, each a system of
    its own.  Water counts log-uniform in [n_min, n_max]; surfactant loading
    uniform in [0, surf_wt_max] by mass; each droplet (with the surfactant
    chains standing out of its surface) rests 3.5 A above a graphite slab whose
    side is three droplet diameters, in the smallest domain of (4 nb2)^3
    boxes of 12 A (nb2 >= 2) that holds the slab and the droplet.

    Returns a list of (System, nb2)."""
    rng = np.random.default_rng(seed)
    m_w = float(sum(ff.EMW))
    m_s = float(sum(ff.EMS))
    rs_A = ff.RS * ff.SIGMA
    tmpl = surfactant_template()
    l_A = float(np.linalg.norm(tmpl[17] - tmpl[0])) * ff.SIGMA  # surfactant length
    out = []
    for k in range(n):
        n_w = int(round(math.exp(rng.uniform(math.log(n_min), math.log(n_max)))))
        wt = rng.uniform(0.0, surf_wt_max)
        n_s = int(wt * n_w * m_w / ((1.0 - wt) * m_s))
        r_A = (3.0 * n_w / WATER_DENSITY_A3 / (4.0 * math.pi)) ** (1.0 / 3.0)
        # outer radius: droplet() anchors site 1 at max(r - 0.65 l, 0.75 r)
        # from the centre and points the chain outward
        r_out = r_A if n_s == 0 else max(r_A + 0.35 * l_A, 0.75 * r_A + l_A) + 1.0
        half = 3.0 * r_A  # slab side: three diameters
        need = max(2.0 * half, 2.0 * r_out + 7.0 + 3.5 + 8.0)  # slab; droplet between slab and top face
        nb2 = max(2, int(math.ceil(need / (4.0 * rs_A))))
        dom = 4 * nb2 * rs_A
        half = min(half, dom / 2.0 - 0.5)
        z_top_A = -dom / 2.0 + 7.0
        d = droplet(n_w, n_s, center_A=(0.0, 0.0, z_top_A + r_out + 3.5), seed=seed + 101 * k)
        g = graphite_sheet(half_side_A=half, z_top=z_top_A / ff.SIGMA, layers=layers)
        out.append((assemble([d], graphite=g, seed=seed + 101 * k), nb2))
    return out


def r2_system(n_water=9997, n_surf=475, seed=20261209, jitter=0.012):
    """The reference's workload of record (its PARAMETER sizes: 9,997 waters +
    475 surfactants, src/mybox.cmns:59-60) as one droplet centred 12 A above
    the domain centre (surfactants by random sequential placement: 475 are too
    many for evenly spread anchors), over the reference's own two-layer graphite sheet
    (formats.make_graphite, top layer at z = -16 sigma), in the reference
    geometry (12^3 L4 boxes of 12 A, nb2 = 3).  Every site is then displaced
    by a normal deviate of `jitter` sigma per axis (~0.04 A), as in a thermal
    configuration: the template surfactant is all-trans at its angle
    equilibria, where the angle and dihedral forces are differences of nearly
    equal numbers."""
    from . import formats
    d = droplet(n_water, n_surf, center_A=(0.0, 0.0, 12.0), seed=seed, spread='rsa')
    s = assemble([d], graphite=formats.make_graphite(), seed=seed)
    s.pos = s.pos + np.random.default_rng(seed + 1000).normal(0.0, jitter, size=s.pos.shape)
    return s
