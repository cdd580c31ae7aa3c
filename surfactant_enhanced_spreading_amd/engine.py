"""Python front end of libsesmd.so (the HIP engine) over its C ABI.

The method names follow the reference's routine families: rebuild() is the
every-10-steps migration + list build, fmm() the far-field chain
computeOnm4..computeLjk4, forces() the eval*/evalLjk4 force routines,
move_a()/move_b() the move_*_a / move_*_b integrator halves and
init_accel() src/init_accel.f90.  Arrays are (N, 3) in reduced units.

There is no CPU fallback: without the built library or a GPU this module
raises.
"""
import ctypes
import os

import numpy as np

from . import forcefield as ff

PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG, 'libsesmd.so')

NEAR, FAR, INTERNAL, SUBTRACT, VALENCE, DIHEDRAL, DIHEDRAL_BE, WALL, GRAPHITE = \
    1, 2, 4, 8, 16, 32, 64, 128, 256
ALL = 511
ALL_MASK = (1 << 64) - 1
FP64, MIXED = 0, 1
P2M, M2M_43, M2M_32, M2L_2, L2L_23, M2L_3, L2L_34, M2L_4 = range(8)


class ForceField(ctypes.Structure):
    """Mirror of sesmd_forcefield (include/sesmd.h)."""
    _fields_ = [
        ('q', ctypes.c_double * 32), ('mass', ctypes.c_double * 32),
        ('lj_c', ctypes.c_double * 1024), ('lj_s6', ctypes.c_double * 1024),
        ('gr_c', ctypes.c_double * 32), ('gr_s6', ctypes.c_double * 32),
        ('dsqw', ctypes.c_double * 3), ('dsqs', ctypes.c_double * 28),
        ('ibond', ctypes.c_int * 28), ('jbond', ctypes.c_int * 28),
        ('th', ctypes.c_double * 36), ('rkth', ctypes.c_double * 36),
        ('ang_i', ctypes.c_int * 36), ('ang_j', ctypes.c_int * 36), ('ang_k', ctypes.c_int * 36),
        ('dva', ctypes.c_double * 30), ('dvb', ctypes.c_double * 30), ('dvc', ctypes.c_double * 30),
        ('dih_i', ctypes.c_int * 30), ('dih_j', ctypes.c_int * 30),
        ('dih_k', ctypes.c_int * 30), ('dih_l', ctypes.c_int * 30),
        ('n_internal', ctypes.c_int),
        ('int_j', ctypes.c_int * 320), ('int_k', ctypes.c_int * 320),
        ('rifact', ctypes.c_double * 33),
        ('xpi', ctypes.c_double), ('rc', ctypes.c_double), ('rs', ctypes.c_double),
        ('dt', ctypes.c_double), ('dtsq2', ctypes.c_double),
        ('clamp_opls', ctypes.c_double), ('clamp_be', ctypes.c_double),
    ]


def reference_forcefield():
    """The reference's force field (src/init_arrays.f90) in the 32-type layout."""
    T = ff.type_tables()
    s = ForceField()
    s.q[:] = list(T['q'])
    s.mass[:] = list(T['mass'])
    s.lj_c[:] = list(T['lj_c'].reshape(-1))
    s.lj_s6[:] = list(T['lj_s6'].reshape(-1))
    s.gr_c[:] = list(T['gr_c'])
    s.gr_s6[:] = list(T['gr_s6'])
    s.dsqw[:] = ff.DSQW
    s.dsqs[:] = ff.DSQS
    s.ibond[:] = ff.IBOND
    s.jbond[:] = ff.JBOND
    s.th[:] = ff.TH
    s.rkth[:] = ff.RKTH
    s.ang_i[:], s.ang_j[:], s.ang_k[:] = ff.ANG_I, ff.ANG_J, ff.ANG_K
    s.dva[:], s.dvb[:], s.dvc[:] = ff.DVA, ff.DVB, ff.DVC
    s.dih_i[:], s.dih_j[:], s.dih_k[:], s.dih_l[:] = ff.DIH_I, ff.DIH_J, ff.DIH_K, ff.DIH_L
    pairs = ff.internal_pairs()
    s.n_internal = len(pairs)
    s.int_j[:len(pairs)] = [p[0] for p in pairs]
    s.int_k[:len(pairs)] = [p[1] for p in pairs]
    s.rifact[:] = ff.RIFACT
    s.xpi, s.rc, s.rs, s.dt, s.dtsq2 = ff.XPI, ff.RC, ff.RS, ff.DT, ff.DTSQ2
    s.clamp_opls, s.clamp_be = ff.dihedral_clamp(1), ff.dihedral_clamp(30)
    return s


_lib = None


def lib():
    """Load libsesmd.so; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError('libsesmd.so not built: run '
                              '`python -m surfactant_enhanced_spreading_amd.csrc.build`')
        L = ctypes.CDLL(LIB_PATH)
        vp, dp, ip = ctypes.c_void_p, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_int)
        L.sesmd_last_error.restype = ctypes.c_char_p
        L.sesmd_create.argtypes = [ctypes.POINTER(vp)] + [ctypes.c_int] * 4 + \
            [ctypes.POINTER(ForceField), ctypes.c_int]
        L.sesmd_destroy.argtypes = [vp]
        L.sesmd_set_state.argtypes = [vp, dp, dp, dp, ctypes.c_double]
        L.sesmd_get_state.argtypes = [vp, dp, dp, dp, dp]
        L.sesmd_set_nh.argtypes = [vp, ctypes.c_double, ctypes.c_double, ctypes.c_double]
        L.sesmd_set_graphite.argtypes = [vp, ctypes.c_int, dp]
        L.sesmd_rebuild.argtypes = [vp]
        L.sesmd_fmm.argtypes = [vp]
        L.sesmd_forces.argtypes = [vp, ctypes.c_uint, ctypes.c_uint64]
        L.sesmd_get_forces.argtypes = [vp, dp]
        L.sesmd_set_forces.argtypes = [vp, dp]
        L.sesmd_move_a.argtypes = [vp, dp]
        L.sesmd_move_b.argtypes = [vp]
        L.sesmd_init_accel.argtypes = [vp]
        L.sesmd_steps.argtypes = [vp, ctypes.c_int, ctypes.c_int, ip]
        L.sesmd_get_classes.argtypes = [vp, ip, ip, ip, ip, ctypes.POINTER(ctypes.c_uint)]
        L.sesmd_set_classes.argtypes = [vp, ip, ip]
        L.sesmd_get_expansions.argtypes = [vp] + [dp] * 6
        L.sesmd_get_counters.argtypes = [vp, ip]
        L.sesmd_device_arrays.argtypes = [vp] + [ctypes.POINTER(vp)] * 4
        L.sesmd_stream.argtypes = [vp]
        L.sesmd_stream.restype = vp
        L.sesmd_near_only.argtypes = [vp]
        L.sesmd_fmm_stage.argtypes = [vp, ctypes.c_int]
        L.sesmd_fmm_stage_enqueue.argtypes = [vp, ctypes.c_int]
        L.sesmd_pair_counts.argtypes = [vp, dp, dp]
        L.sesmd_get_thermostat.argtypes = [vp, dp, dp]
        L.sesmd_set_thermostat.argtypes = [vp, ctypes.c_double, ctypes.c_double]
        L.sesmd_get_flags.argtypes = [vp, ip, ctypes.c_int]
        L.sesmd_energies.argtypes = [vp, dp]
        L.sesmd_ensemble_create.argtypes = [ctypes.POINTER(vp), ctypes.POINTER(vp), ctypes.c_int]
        L.sesmd_ensemble_destroy.argtypes = [vp]
        L.sesmd_ensemble_set_shape.argtypes = [vp, ctypes.c_int, ctypes.c_int, ctypes.c_int]
        L.sesmd_ensemble_get_shape.argtypes = [vp, ip]
        L.sesmd_ensemble_tune.argtypes = [vp, ctypes.c_int, ip, ctypes.POINTER(ctypes.c_float)]
        L.sesmd_ensemble_near_only.argtypes = [vp]
        L.sesmd_ensemble_stream.argtypes = [vp]
        L.sesmd_ensemble_stream.restype = vp
        L.sesmd_ensemble_batched.argtypes = [vp]
        L.sesmd_ensemble_steps.argtypes = [vp, ctypes.c_int, ctypes.c_int, ip]
        _lib = L
    return _lib


def _p(a, t=ctypes.c_double):
    return a.ctypes.data_as(ctypes.POINTER(t))


# sesmd_energies order (SESMD_E_* in include/sesmd.h)
ENERGY_NAMES = ('near_coul', 'near_lj', 'far', 'subtract', 'internal_coul', 'internal_lj',
                'angle', 'dihedral', 'dihedral_be', 'wall', 'graphite', 'potential', 'kinetic',
                'thermostat', 'total', 'virial', 'temperature')


class SesmdError(RuntimeError):
    pass


def _ck(rc):
    if rc != 0:
        raise SesmdError(lib().sesmd_last_error().decode())


class Engine:
    """One system resident on the GPU."""

    def __init__(self, sys_, nb2=3, p=ff.P_DEFAULT, precision=FP64, tr=None, eta=0.0,
                 forcefield=None):
        self.L = lib()
        self._ff = forcefield if forcefield is not None else reference_forcefield()
        self.n_water, self.n_surf = sys_.n_water, sys_.n_surf
        self.n = sys_.n_sites
        self.nb2, self.p = nb2, p
        self.nb = 4 * nb2
        self.lp = (p + 1) ** 2
        h = ctypes.c_void_p()
        _ck(self.L.sesmd_create(ctypes.byref(h), self.n_water, self.n_surf, nb2, p,
                                ctypes.byref(self._ff), precision))
        self.h = h
        acc = sys_.acc if sys_.acc is not None else np.zeros_like(sys_.pos)
        self.set_state(sys_.pos, sys_.vel, acc, eta)
        tr = ff.TR_DEFAULT if tr is None else tr
        npart = 3 * self.n_water + 29 * self.n_surf
        nc = 3 * self.n_water + 28 * self.n_surf
        df = 3 * npart - nc
        self.df = df
        _ck(self.L.sesmd_set_nh(self.h, tr, ff.DT / df / tr / (ff.RLXTIME * ff.RLXTIME),
                                -2.0 * df * tr))
        g = np.ascontiguousarray(sys_.graphite, dtype=np.float64).reshape(-1, 3)
        _ck(self.L.sesmd_set_graphite(self.h, len(g), _p(g)))

    def close(self):
        if getattr(self, 'h', None):
            self.L.sesmd_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # state ------------------------------------------------------------------
    def set_state(self, r, v, a, eta):
        r = np.ascontiguousarray(r, dtype=np.float64)
        v = np.ascontiguousarray(v, dtype=np.float64)
        a = np.ascontiguousarray(a, dtype=np.float64)
        _ck(self.L.sesmd_set_state(self.h, _p(r), _p(v), _p(a), float(eta)))

    def state(self):
        r = np.zeros((self.n, 3))
        v = np.zeros((self.n, 3))
        a = np.zeros((self.n, 3))
        eta = ctypes.c_double()
        _ck(self.L.sesmd_get_state(self.h, _p(r), _p(v), _p(a), ctypes.byref(eta)))
        return r, v, a, eta.value

    def thermostat(self):
        """(eta, xi): the Nose-Hoover friction and its time integral."""
        eta, xi = ctypes.c_double(), ctypes.c_double()
        _ck(self.L.sesmd_get_thermostat(self.h, ctypes.byref(eta), ctypes.byref(xi)))
        return eta.value, xi.value

    def set_thermostat(self, eta, xi):
        _ck(self.L.sesmd_set_thermostat(self.h, float(eta), float(xi)))

    def set_classes(self, cls, anc):
        """Molecule classes (0 W1, 1 W2, 2 surfactant) and storage boxes that
        every rebuild keeps instead of classifying (sesmd_set_classes); None
        lets the rebuilds classify again, from the classes held now."""
        if cls is None:
            _ck(self.L.sesmd_set_classes(self.h, None, None))
            return
        cls = np.ascontiguousarray(cls, dtype=np.int32)
        anc = np.ascontiguousarray(anc, dtype=np.int32)
        _ck(self.L.sesmd_set_classes(self.h, _p(cls, ctypes.c_int), _p(anc, ctypes.c_int)))

    def checkpoint(self, path, kb):
        """Write the whole dynamical state (positions, v dt, F dt^2/2m, eta, xi),
        the molecule classes and storage boxes of the last rebuild (part A of
        the next step groups its kinetic sums by class) and the step counter
        kb to an .npz file."""
        r, v, a, _ = self.state()
        eta, xi = self.thermostat()
        c = self.classes()
        np.savez(path, r=r, v=v, a=a, eta=eta, xi=xi, kb=kb, n_water=self.n_water, n_surf=self.n_surf,
                 nb2=self.nb2, p=self.p, cls=c['cls'], anc=c['anc'])

    def restore(self, path):
        """Load a checkpoint written by checkpoint() into this context (same
        system size and geometry), rebuild, and return its step counter."""
        d = np.load(path, allow_pickle=False)
        if (int(d['n_water']), int(d['n_surf']), int(d['nb2']), int(d['p'])) != \
                (self.n_water, self.n_surf, self.nb2, self.p):
            raise SesmdError('checkpoint is of another system or geometry')
        self.set_state(d['r'], d['v'], d['a'], float(d['eta']))
        self.set_thermostat(float(d['eta']), float(d['xi']))
        if 'cls' in d:
            # lists from the restored positions with the checkpointed classes,
            # which part A of the next step reads; the next rebuild classifies
            # again
            self.set_classes(d['cls'], d['anc'])
            self.rebuild()
            self.set_classes(None, None)
        else:
            self.rebuild()
        return int(d['kb'])

    def flags(self):
        """Device error counters: SHAKE/RATTLE non-convergence, steps with
        non-finite kinetic sums, molecules filed outside the cube, mixed
        near-field items whose fixed-point accumulation could leave its range."""
        f = (ctypes.c_int * 4)()
        _ck(self.L.sesmd_get_flags(self.h, f, 4))
        return dict(shake=f[0], nonfinite=f[1], outside=f[2], fixed_range=f[3])

    # reference routine families ---------------------------------------------
    def rebuild(self):
        _ck(self.L.sesmd_rebuild(self.h))

    def fmm(self):
        _ck(self.L.sesmd_fmm(self.h))

    def forces(self, what=ALL, mask=ALL_MASK, fetch=True):
        _ck(self.L.sesmd_forces(self.h, what, mask))
        if not fetch:
            return None
        f = np.zeros((self.n, 3))
        _ck(self.L.sesmd_get_forces(self.h, _p(f)))
        return f

    def set_forces(self, f):
        f = np.ascontiguousarray(f, dtype=np.float64)
        _ck(self.L.sesmd_set_forces(self.h, _p(f)))

    def move_a(self):
        s = np.zeros(6)
        _ck(self.L.sesmd_move_a(self.h, _p(s)))
        return s

    def move_b(self):
        _ck(self.L.sesmd_move_b(self.h))

    def init_accel(self):
        _ck(self.L.sesmd_init_accel(self.h))

    def steps(self, nsteps, ifar=10, kb=0):
        k = ctypes.c_int(kb)
        _ck(self.L.sesmd_steps(self.h, nsteps, ifar, ctypes.byref(k)))
        return k.value

    def bootstrap(self):
        """The start-up force evaluation of src/mybox.f90:65-128."""
        self.rebuild()
        self.fmm()
        self.forces(fetch=False)
        self.init_accel()

    # inspection -------------------------------------------------------------
    def classes(self):
        nm = self.n_water + self.n_surf
        cls = np.zeros(nm, np.int32)
        anc = np.zeros(nm, np.int32)
        cbox = np.zeros(self.n, np.int32)
        rk = np.zeros(self.n, np.int32)
        isub = np.zeros(self.n, np.uint32)
        _ck(self.L.sesmd_get_classes(self.h, _p(cls, ctypes.c_int), _p(anc, ctypes.c_int),
                                     _p(cbox, ctypes.c_int), _p(rk, ctypes.c_int),
                                     _p(isub, ctypes.c_uint)))
        return dict(cls=cls, anc=anc, cbox=cbox, rkind=rk, isub=isub)

    def expansions(self):
        nb3, n33, n2, lp = self.nb ** 3, (self.nb // 2) ** 3, self.nb2 ** 3, self.lp
        arrs = [np.zeros((nb3, lp), np.complex128), np.zeros((nb3, lp), np.complex128),
                np.zeros((n33, lp), np.complex128), np.zeros((n33, lp), np.complex128),
                np.zeros((n2, lp), np.complex128), np.zeros((n2, lp), np.complex128)]
        _ck(self.L.sesmd_get_expansions(self.h, *[_p(a) for a in arrs]))
        return dict(zip(('O4', 'L4', 'O3', 'L3', 'O2', 'L2'), arrs))

    def shake_failures(self):
        k = ctypes.c_int()
        _ck(self.L.sesmd_get_counters(self.h, ctypes.byref(k)))
        return k.value

    def pair_counts(self):
        a, b = ctypes.c_double(), ctypes.c_double()
        _ck(self.L.sesmd_pair_counts(self.h, ctypes.byref(a), ctypes.byref(b)))
        return a.value, b.value

    def stream(self):
        return self.L.sesmd_stream(self.h)

    def energies(self):
        """Energy components at the current state (sesmd_energies; FP64)."""
        e = np.zeros(len(ENERGY_NAMES))
        _ck(self.L.sesmd_energies(self.h, _p(e)))
        return dict(zip(ENERGY_NAMES, e))

    def near_only(self):
        _ck(self.L.sesmd_near_only(self.h))

    def fmm_stage(self, stage, wait=True):
        """One far-field stage (SESMD_P2M .. SESMD_M2L_4); wait=False only
        enqueues it on the context's stream."""
        if wait:
            _ck(self.L.sesmd_fmm_stage(self.h, stage))
        else:
            _ck(self.L.sesmd_fmm_stage_enqueue(self.h, stage))

    def temperature(self):
        """Kinetic temperature sum m v^2 / dt^2 / df (src/compute_temperature.f90:240-244)."""
        r, v, a, eta = self.state()
        m = np.concatenate([np.tile(ff.EMW, self.n_water), np.tile(ff.EMS, self.n_surf)])
        return float((m[:, None] * v * v).sum() / ff.DT ** 2 / self.df)


class Ensemble:
    """Independent systems advanced together (sesmd_ensemble_*): every
    member's step in one captured graph, the near field of all members in one
    persistent launch.  Members are Engine objects of the same precision,
    built (bootstrap() done); they must stay open while the ensemble lives,
    and each keeps its own state (state(), forces() ... read it as usual)."""

    def __init__(self, members):
        self.L = lib()
        self.members = list(members)
        arr = (ctypes.c_void_p * len(self.members))(*[m.h.value for m in self.members])
        h = ctypes.c_void_p()
        _ck(self.L.sesmd_ensemble_create(ctypes.byref(h), arr, len(self.members)))
        self.h = h

    def close(self):
        if getattr(self, 'h', None):
            self.L.sesmd_ensemble_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def steps(self, nsteps, ifar=10, kb=0):
        k = ctypes.c_int(kb)
        _ck(self.L.sesmd_ensemble_steps(self.h, nsteps, ifar, ctypes.byref(k)))
        return k.value

    def tune(self, reps=3):
        """The autotuner: returns ((waves, chunk, grid), ms per launch) of the
        fastest near-field launch shape, which the ensemble then keeps."""
        sh = (ctypes.c_int * 3)()
        ms = ctypes.c_float()
        _ck(self.L.sesmd_ensemble_tune(self.h, reps, sh, ctypes.byref(ms)))
        return tuple(sh), ms.value

    def set_shape(self, waves, chunk, grid):
        _ck(self.L.sesmd_ensemble_set_shape(self.h, waves, chunk, grid))

    def shape(self):
        sh = (ctypes.c_int * 3)()
        _ck(self.L.sesmd_ensemble_get_shape(self.h, sh))
        return tuple(sh)

    def batched(self):
        """True when the members' integrator and site-term launches run as
        one launch each per step."""
        return self.L.sesmd_ensemble_batched(self.h) == 1

    def near_only(self):
        _ck(self.L.sesmd_ensemble_near_only(self.h))

    def stream(self):
        return self.L.sesmd_ensemble_stream(self.h)
