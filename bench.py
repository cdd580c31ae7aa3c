#!/usr/bin/env python3
"""Throughput of the per-timestep force + integration path on one MI355X.

Headline workload: one droplet of 20,000 rigid SPC/E waters carrying 200
29-site M(D'E4OH)M surfactants (65,800 sites) in a 16^3-box domain (nb2 = 4:
4,096 L4 boxes of 12 A, a 192 A cube), expansion order p = 8, far field and
lists refreshed every 10 steps, 2 fs, Nose-Hoover at 298 K.  A step is the
reference's full step (src/mybox.f90:130-227); the timed region covers K
consecutive steps, rebuild steps included, with all inputs resident in HBM.
`value` is the FP64 path (the drop-in's precision); the "mixed" object reports
the mixed-precision mode (FP32 pair arithmetic) on the same workload, its
speedup and its RMS force error against FP64, at the start and after the
mixed drift run.  "reference_geometry" repeats the timing in the reference's
own geometry (12^3 boxes, p = 4).

Prints one JSON line.  The library and torch ship different HSA runtimes, so
timing uses HIP events of the library's own runtime (/opt/rocm), not torch.
"""
import argparse
import ctypes
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "MD steps/s (ns/day) 20k water+surfactant: mixed vs FP64 GPU at RMS F err≤1e-5"
FP64_PEAK_TFLOPS = 78.6
FP32_PEAK_TFLOPS = 157.3
BF16_PEAK_TFLOPS = 2500.0  # dense bf16 matrix peak
HBM_PEAK_GBS = 8000.0


class HipTimer:
    """hipEvent pairs on a given stream (the runtime libsesmd.so uses)."""

    def __init__(self):
        self.hip = ctypes.CDLL('/opt/rocm/lib/libamdhip64.so')
        self.hip.hipEventCreate.argtypes = [ctypes.POINTER(ctypes.c_void_p)]
        self.hip.hipEventRecord.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
        self.hip.hipEventSynchronize.argtypes = [ctypes.c_void_p]
        self.hip.hipEventElapsedTime.argtypes = [ctypes.POINTER(ctypes.c_float),
                                                 ctypes.c_void_p, ctypes.c_void_p]
        self.hip.hipStreamSynchronize.argtypes = [ctypes.c_void_p]
        self.hip.hipDeviceSynchronize.argtypes = []

    def event(self):
        e = ctypes.c_void_p()
        assert self.hip.hipEventCreate(ctypes.byref(e)) == 0
        return e

    def record(self, ev, stream):
        assert self.hip.hipEventRecord(ev, ctypes.c_void_p(stream)) == 0

    def elapsed_ms(self, e0, e1):
        assert self.hip.hipEventSynchronize(e1) == 0
        ms = ctypes.c_float()
        assert self.hip.hipEventElapsedTime(ctypes.byref(ms), e0, e1) == 0
        return ms.value

    def sync(self):
        assert self.hip.hipDeviceSynchronize() == 0


# The two workloads: the headline (20k + 200 droplet, 16^3 boxes, p = 8) and
# the same droplet in the reference's own geometry (12^3 boxes, p = 4).
CONFIGS = {
    'c2': dict(nb2=4, p=8, workload="20k SPC/E water + 200 M(D'E4OH)M surfactant droplet, "
                                    "16^3 L4 boxes of 12 A (4,096 boxes), p=8, ifar=10, 2 fs"),
    'ref': dict(nb2=3, p=4, workload="the same droplet in the reference geometry "
                                     "(12^3 L4 boxes of 12 A, p=4, ifar=10, 2 fs)"),
}


def make_system(name):
    from surfactant_enhanced_spreading_amd import builder
    return builder.config1_system() if name == 'c2' else builder.bench_system()


def cpu_baseline(sys_, nb2, p, seconds=20.0):
    """The C restatement (oracle) on the host cores: one rebuild step (classes,
    far field, forces) then regular steps until ~`seconds` have passed."""
    sys.path.insert(0, os.path.join(ROOT, 'oracle'))
    import oracle as orc
    threads = min(16, os.cpu_count() or 1)
    orc.lib().orc_set_threads(threads)
    o = orc.Oracle(sys_, nb2=nb2, p=p)
    o.rebuild()
    o.fmm()
    o.init_accel(o.forces())
    kb = 9  # next step (10) is a rebuild step, as in the GPU timed region
    t0 = time.perf_counter()
    n = 0
    while True:
        kb = o.steps(1, 10, kb)
        n += 1
        el = time.perf_counter() - t0
        if el > seconds or n >= 50:
            break
    return dict(value=n / el, unit='steps/s', cores=threads, kind='port',
                sample='%d full steps (1 rebuild+far-field step, %d regular) of the same '
                       '%d-site system at p=%d, C restatement with OpenMP'
                       % (n, n - 1, sys_.n_sites, p))


def run(sys_, cfg, prec, steps, warmup, timer):
    """Bootstrap, warm up, then time `steps` full steps and the near-field
    kernel alone; returns the figures and the engine (still open).  The first
    steps() call captures the step graphs before it runs any step, so no
    capture falls inside the timed region."""
    from surfactant_enhanced_spreading_amd import engine
    e = engine.Engine(sys_, nb2=cfg['nb2'], p=cfg['p'], precision=prec)
    e.bootstrap()
    stream = e.stream()
    kb = e.steps(warmup, 10, 0)
    timer.sync()
    e0, e1 = timer.event(), timer.event()
    t0 = time.perf_counter()
    timer.record(e0, stream)
    kb = e.steps(steps, 10, kb)
    timer.record(e1, stream)
    ms = timer.elapsed_ms(e0, e1)
    timer.sync()
    wall = time.perf_counter() - t0
    # dominant kernel: the near field, launched alone on the same stream
    e.forces(fetch=False)
    reps = 50
    timer.sync()
    timer.record(e0, stream)
    for _ in range(reps):
        e.near_only()
    timer.record(e1, stream)
    near_ms = timer.elapsed_ms(e0, e1) / reps
    timer.sync()
    ms_step = ms / steps
    return dict(ms_step=ms_step, steps_s=1000.0 / ms_step, wall=wall, near_ms=near_ms), e


def m2l_translations(O4, nb):
    """(target, source) L4 translations with a non-empty source: for every L4
    box, the children of its parent's 27 neighbours less its own 27 neighbours
    (src/computeLjk4.f90:58-170), counted where the source holds charge."""
    import numpy as np
    ne = (np.abs(O4).sum(axis=1) != 0).reshape(nb, nb, nb)  # [z][y][x]
    total = 0
    idx = np.arange(nb)
    for dz in range(-3, 4):
        for dy in range(-3, 4):
            for dx in range(-3, 4):
                if max(abs(dx), abs(dy), abs(dz)) <= 1:
                    continue
                # targets t whose parent neighbours the source's parent
                ok = []
                for d in (dx, dy, dz):
                    t = idx
                    s = t + d
                    ok.append((s >= 0) & (s < nb) & (np.abs(s // 2 - t // 2) <= 1))
                okx, oky, okz = ok
                m = okz[:, None, None] & oky[None, :, None] & okx[None, None, :]
                zz, yy, xx = np.nonzero(m)
                total += int(ne[zz + dz, yy + dy, xx + dx].sum())
    return total


def m2l_timing(e, timer, reps=20, mixed=False):
    """The L4 M2L (computeLjk4) alone, `reps` launches on the context's stream
    after a rebuild and far field at the current state.  Work per translation:
    the real-linear map of the 2 lk real inputs Re, Im O(n, m >= 0) to the 2 lk
    real outputs, 2 (2 lk)^2 flop (lk = (p+1)(p+2)/2), which is what the
    products compute (the m < 0 half of an expansion is the conjugate of the
    m > 0 half); the reference's complex sum over all (p+1)^2 inputs,
    8 lk (p+1)^2 flop, is reported beside it.  FP64 context: against the FP64
    matrix peak.  Mixed context (split-BF16x3 products): the six bf16 products
    per translation against the dense bf16 matrix peak."""
    from surfactant_enhanced_spreading_amd import engine
    e.rebuild()
    e.fmm()
    T = m2l_translations(e.expansions()['O4'], e.nb)
    stream = e.stream()
    e0, e1 = timer.event(), timer.event()
    timer.sync()
    timer.record(e0, stream)
    for _ in range(reps):
        e.fmm_stage(engine.M2L_4, wait=False)
    timer.record(e1, stream)
    ms = timer.elapsed_ms(e0, e1) / reps
    timer.sync()
    p = e.p
    lk = (p + 1) * (p + 2) // 2
    flop = 2.0 * (2 * lk) ** 2 * T
    ach = flop / (ms * 1e-3) / 1e12
    ref = 8.0 * lk * (p + 1) ** 2 * T / (ms * 1e-3) / 1e12
    if mixed:
        return dict(kernel='M2L at L4 (k_m2l_count, k_m2l_scan, k_m2l_fill, k_m2l_split, k_m2l_gemm_bf3, '
                           'k_m2l_accum)', translations=T, ms=round(ms, 4),
                    achieved_tflops=round(ach, 3), reference_count_tflops=round(ref, 3),
                    bf16_products_tflops=round(6 * ach, 3), peak_bf16_tflops=BF16_PEAK_TFLOPS,
                    frac_bf16_mfma_peak=round(6 * ach / BF16_PEAK_TFLOPS, 4))
    return dict(kernel='M2L at L4 (k_m2l_count, k_m2l_scan, k_m2l_fill, k_m2l_gemm, k_m2l_accum)', translations=T, ms=round(ms, 4),
                achieved_tflops=round(ach, 3), reference_count_tflops=round(ref, 3), peak_tflops=FP64_PEAK_TFLOPS,
                frac_fp64_mfma_peak=round(ach / FP64_PEAK_TFLOPS, 4))


def energy_drift(sys_, cfg, prec, nsteps, every=500, settle=100, ifar=10, keep=False):
    """Extended (Nose-Hoover) energy over `nsteps` steps after `settle` steps
    (the first steps project the start-up velocities onto the constraints):
    drift and largest excursion relative to the kinetic energy, energies
    evaluated in FP64 (sesmd_energies) every `every` steps.  ifar: steps
    between list / far-field refreshes (the reference's 10).  keep: also
    return the final state."""
    from surfactant_enhanced_spreading_amd import engine
    e = engine.Engine(sys_, nb2=cfg['nb2'], p=cfg['p'], precision=prec)
    try:
        e.bootstrap()
        kb = e.steps(settle, ifar, 0)
        e0 = e.energies()
        hs = [e0['total']]
        done = 0
        while done < nsteps:
            k = min(every, nsteps - done)
            kb = e.steps(k, ifar, kb)
            done += k
            hs.append(e.energies()['total'])
        e1 = e.energies()
        final = e.state() if keep else None
    finally:
        e.close()
    k0 = e0['kinetic']
    ns = nsteps * 2.0e-6  # 2 fs steps
    out = dict(steps=nsteps, ifar=ifar, kinetic0=round(k0, 3), total0=round(hs[0], 4),
               total_end=round(hs[-1], 4),
               drift_rel_kinetic=(hs[-1] - hs[0]) / k0,
               drift_rel_kinetic_per_ns=(hs[-1] - hs[0]) / k0 / ns,
               max_excursion_rel_kinetic=max(abs(h - hs[0]) for h in hs) / k0,
               temperature0=round(e0['temperature'], 4), temperature_end=round(e1['temperature'], 4))
    return (out, final) if keep else out


def rms_force_error(sys_, cfg, state=None):
    """||F_mixed - F_fp64|| / ||F_fp64|| over all sites and components, total
    force after a fresh rebuild and far field, at the system's configuration
    or at `state` (positions, velocities, accelerations, eta)."""
    import numpy as np
    from surfactant_enhanced_spreading_amd import engine
    f = []
    for prec in (engine.FP64, engine.MIXED):
        e = engine.Engine(sys_, nb2=cfg['nb2'], p=cfg['p'], precision=prec)
        if state is not None:
            e.set_state(*state)
        e.rebuild()
        e.fmm()
        f.append(e.forces())
        e.close()
    return float(np.sqrt(((f[1] - f[0]) ** 2).sum() / (f[0] ** 2).sum()))


def ensemble_leg(timer, steps, warmup, n=64):
    """The batched-ensemble workload: n droplets (builder.ensemble_systems),
    mixed precision, p = 8, one context each, advanced by one ensemble (one
    graph per step, the near field of all droplets in one persistent launch
    whose shape the autotuner picks).  Times `steps` ensemble steps after
    `warmup` (whole ten-step cycles; graphs captured before the timed region)
    and the ensemble near-field launch alone; the FP32 rate counts 21 flops
    per pair and 9 more per LJ pair, as the headline roofline does."""
    from surfactant_enhanced_spreading_amd import builder, engine
    t0 = time.perf_counter()
    systems = builder.ensemble_systems(n=n)
    t_build = time.perf_counter() - t0
    t0 = time.perf_counter()
    members = []
    for s, nb2 in systems:
        e = engine.Engine(s, nb2=nb2, p=8, precision=engine.MIXED)
        e.bootstrap()
        members.append(e)
    t_create = time.perf_counter() - t0
    print('bench: ensemble of %d built (%.1f s) and created (%.1f s)' % (n, t_build, t_create),
          file=sys.stderr, flush=True)
    pairs = lj = 0.0
    for e in members:
        a, b = e.pair_counts()
        pairs += a
        lj += b
    ens = engine.Ensemble(members)
    shape, tune_ms = ens.tune(reps=3)
    stream = ens.stream()
    kb = ens.steps(warmup, 10, 9)  # kb = 9: the first step is a rebuild, every run whole cycles
    timer.sync()
    e0, e1 = timer.event(), timer.event()
    timer.record(e0, stream)
    kb = ens.steps(steps, 10, kb)
    timer.record(e1, stream)
    ms = timer.elapsed_ms(e0, e1) / steps
    timer.sync()
    reps = 10
    timer.record(e0, stream)
    for _ in range(reps):
        ens.near_only()
    timer.record(e1, stream)
    near_ms = timer.elapsed_ms(e0, e1) / reps
    timer.sync()
    flags = [e.flags() for e in members]
    ens.close()
    sites = sum(s.n_sites for s, _ in systems)
    for e in members:
        e.close()
    ach = (21.0 * pairs + 9.0 * lj) / (near_ms * 1e-3) / 1e12
    return dict(workload='%d droplets, waters log-uniform 500-30,000, 0-5 wt%% surfactant, graphite slab '
                         '3 diameters wide, domains of (4 nb2)^3 boxes of 12 A by size, p=8, mixed '
                         'precision, ifar=10, 2 fs' % n,
                droplets=n, sites=sites, carbons=int(sum(len(s.graphite) for s, _ in systems)),
                nb2_counts={str(k): sum(1 for _, b in systems if b == k) for k in sorted({b for _, b in systems})},
                steps=steps, warmup=warmup, ms_per_step=round(ms, 4),
                steps_per_s=round(1000.0 / ms, 2),
                droplet_ns_per_day=round(1000.0 / ms * n * 2.0 * 0.0864, 2),
                near_kernel='k_near_rows_ens (persistent, %d waves/workgroup, %d receivers/wave, %d workgroups)'
                            % shape,
                near_kernel_ms=round(near_ms, 4), tune_ms=round(tune_ms, 4),
                pairs=pairs, lj_pairs=lj, near_achieved_tflops=round(ach, 3),
                near_frac_fp32_peak=round(ach / FP32_PEAK_TFLOPS, 4),
                near_frac_fp32_unpacked_peak=round(ach / FP64_PEAK_TFLOPS, 4),
                shake_failures=sum(f['shake'] for f in flags),
                flags_nonzero=sum(1 for f in flags if any(f.values())),
                build_s=round(t_build, 1), create_s=round(t_create, 1))


def reference_leg(timer, seconds, steps, warmup):
    """The reference's workload of record (9,997 waters + 475 surfactants, its
    PARAMETER sizes, src/mybox.cmns:59-60) as the data_cont2.dat its program
    reads: the GPU engine on it in the reference geometry (12^3 boxes, p = 4),
    FP64 and mixed, and beside it the reference program itself
    (oracle/_ref/mybox_ref, 27 MPI ranks on this host's cores, about
    `seconds` of stepping; median step time from step 21 on, regular and
    rebuild + far-field steps apart)."""
    import tempfile
    sys.path.insert(0, os.path.join(ROOT, 'oracle'))
    import run_reference as rr
    from surfactant_enhanced_spreading_amd import builder, engine, formats
    s0 = builder.r2_system()
    with tempfile.TemporaryDirectory() as d:
        path = os.path.join(d, 'data_cont2.dat')
        formats.write_config(path, s0)
        s = formats.read_config(path, s0.n_water, s0.n_surf)
    s.graphite = s0.graphite
    cfg = dict(nb2=3, p=4)
    out = dict(workload='9,997 SPC/E waters + 475 M(D\'E4OH)M surfactants (the reference\'s PARAMETER '
                        'sizes) as one droplet over the reference\'s own graphite sheet, 12^3 L4 boxes '
                        'of 12 A, p=4, ifar=10, 2 fs', sites=s.n_sites)
    for prec, key in ((engine.FP64, 'gpu_fp64'), (engine.MIXED, 'gpu_mixed')):
        r, e = run(s, cfg, prec, steps, warmup, timer)
        out[key] = dict(steps_per_s=round(r['steps_s'], 2), ms_per_step=round(r['ms_step'], 4),
                        shake_failures=e.shake_failures())
        e.close()
    if rr.available():
        res = rr.run(s, seconds)
        sm = rr.summary(res, skip=20)
        ref = dict(kind='reference', program='oracle/_ref/mybox_ref (src/mybox.f90 and its routines, '
                                             'amdflang -O2, MPICH 3.3.2)',
                   ranks=sm['ranks'], cores=sm['cores'], steps_timed=sm['steps'],
                   startup_s=round(res['startup_s'], 2) if res['startup_s'] else None,
                   shake_errors=res['shake_errors'])
        for k in ('regular_median_s', 'far_field_median_s', 'cycle_mean_s_per_step'):
            if k in sm:
                ref[k] = round(sm[k], 4)
        if 'cycle_mean_s_per_step' in sm:
            ref['steps_per_s'] = round(1.0 / sm['cycle_mean_s_per_step'], 4)
            out['gpu_fp64_over_reference'] = round(out['gpu_fp64']['steps_per_s'] * sm['cycle_mean_s_per_step'], 1)
        out['reference_cpu'] = ref
    else:
        out['reference_cpu'] = None
    return out


TRAFFIC_PROFILE = os.path.join(ROOT, 'profiles', 'k_near_rows_counters_c2_fp64.md')


def near_traffic():
    """HBM bytes per k_near_rows launch on the headline workload, from the
    committed FETCH_SIZE / WRITE_SIZE counter runs (each scaled by what the
    counter reads for a plain copy of known size); None if that summary is
    absent."""
    try:
        with open(TRAFFIC_PROFILE) as fh:
            for line in fh:
                if line.startswith('traffic_bytes_per_launch:'):
                    return int(line.split(':', 1)[1])
    except OSError:
        pass
    return None


def copy_calibration(nbytes, reps=5):
    """A plain device-to-device copy of `nbytes` (hipMemcpy, the runtime's copy
    kernel), for scaling the FETCH_SIZE / WRITE_SIZE counters of the profiled
    kernels: each copy reads and writes exactly nbytes."""
    hip = ctypes.CDLL('/opt/rocm/lib/libamdhip64.so')
    hip.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    hip.hipMemset.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_size_t]
    hip.hipFree.argtypes = [ctypes.c_void_p]
    a, b = ctypes.c_void_p(), ctypes.c_void_p()
    assert hip.hipMalloc(ctypes.byref(a), nbytes) == 0 and hip.hipMalloc(ctypes.byref(b), nbytes) == 0
    assert hip.hipMemset(a, 1, nbytes) == 0
    for _ in range(reps):
        assert hip.hipMemcpy(b, a, nbytes, 3) == 0  # hipMemcpyDeviceToDevice
    assert hip.hipDeviceSynchronize() == 0
    hip.hipFree(a)
    hip.hipFree(b)
    print(json.dumps(dict(copy_bytes=nbytes, reps=reps)), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument('--gpus', type=int, default=1)
    ap.add_argument('--steps', type=int, default=500)
    ap.add_argument('--warmup', type=int, default=30)
    ap.add_argument('--config', choices=sorted(CONFIGS), default='c2',
                    help='c2: the headline workload; ref: only the reference-geometry workload')
    ap.add_argument('--no-mixed', action='store_true', help='skip the mixed-precision leg')
    ap.add_argument('--no-cpu', action='store_true', help='skip the CPU baseline leg')
    ap.add_argument('--no-secondary', action='store_true', help='skip the reference-geometry leg')
    ap.add_argument('--cpu-seconds', type=float, default=20.0)
    ap.add_argument('--drift-steps', type=int, default=10000,
                    help='steps of the energy-drift runs (FP64 and mixed); 0 skips them')
    ap.add_argument('--ifar-steps', type=int, default=2000,
                    help='steps of each run of the drift-vs-ifar series (ifar 1, 2, 5, 10); 0 skips it')
    ap.add_argument('--no-ensemble', action='store_true', help='skip the batched-ensemble leg')
    ap.add_argument('--no-reference', action='store_true',
                    help='skip the reference-program leg (its workload of record, timed beside the GPU)')
    ap.add_argument('--reference-seconds', type=float, default=40.0)
    ap.add_argument('--reference-only', action='store_true', help='run only the reference-program leg')
    ap.add_argument('--ensemble-only', action='store_true', help='run only the batched-ensemble leg')
    ap.add_argument('--ensemble-steps', type=int, default=50)
    ap.add_argument('--copy-calibrate', type=int, default=0, metavar='BYTES',
                    help='only run a device copy of BYTES (counter calibration) and exit')
    args = ap.parse_args()
    if args.copy_calibrate:
        copy_calibration(args.copy_calibrate)
        return
    if args.gpus != 1:
        print('bench.py measures one GPU: the whole system runs on one MI355X '
              'and there is no multi-GPU path', file=sys.stderr)
        sys.exit(2)

    from surfactant_enhanced_spreading_amd import engine
    name = args.config
    cfg = CONFIGS[name]
    timer = HipTimer()
    if args.ensemble_only:
        print(json.dumps(ensemble_leg(timer, args.ensemble_steps, 10)), flush=True)
        return
    if args.reference_only:
        print(json.dumps(reference_leg(timer, args.reference_seconds, 200, 20)), flush=True)
        return
    sys_ = make_system(name)
    print('bench: %s, %d sites, nb2=%d, p=%d' % (name, sys_.n_sites, cfg['nb2'], cfg['p']),
          file=sys.stderr, flush=True)
    r64, e = run(sys_, cfg, engine.FP64, args.steps, args.warmup, timer)
    pairs, lj_pairs = e.pair_counts()
    shake_failures, temperature = e.shake_failures(), e.temperature()
    m2l = m2l_timing(e, timer)
    e.close()
    steps_s = r64['steps_s']
    flop = 21.0 * pairs + 9.0 * lj_pairs
    achieved = flop / (r64['near_ms'] * 1e-3) / 1e12
    roofline = dict(bound='fp64-valu', achieved=round(achieved, 3), peak=FP64_PEAK_TFLOPS,
                    unit='TFLOP/s', frac=round(achieved / FP64_PEAK_TFLOPS, 4),
                    traffic=near_traffic() if name == 'c2' else None,
                    kernel='k_near_rows', kernel_ms=round(r64['near_ms'], 4),
                    pairs=pairs, lj_pairs=lj_pairs)
    nb = 4 * cfg['nb2']
    out = dict(metric=METRIC, value=round(steps_s, 2), unit='steps/s', n_gpus=1,
               steps=args.steps, warmup=args.warmup, ms_per_step=round(r64['ms_step'], 4),
               higher_is_better=True, scaling='weak', vs_baseline=None, dtype='f64',
               data='synthetic',
               config=dict(workload=cfg['workload'], n_water=sys_.n_water, n_surf=sys_.n_surf,
                           sites=sys_.n_sites, boxes=nb ** 3, p=cfg['p'], ifar=10, dt_fs=2.0,
                           precision='fp64'),
               ns_per_day=round(steps_s * 2.0 * 0.0864, 2),  # dt = 2 fs (src/mybox.cmns:43)
               wall_s=round(r64['wall'], 3), shake_failures=shake_failures,
               temperature=round(temperature, 4), roofline=roofline, m2l=m2l)
    print('bench: fp64 %.1f steps/s' % steps_s, file=sys.stderr, flush=True)
    if not args.no_mixed:
        rm, e = run(sys_, cfg, engine.MIXED, args.steps, args.warmup, timer)
        m2l_mixed = m2l_timing(e, timer, mixed=True)
        e.close()
        ach = flop / (rm['near_ms'] * 1e-3) / 1e12
        out['mixed'] = dict(steps_per_s=round(rm['steps_s'], 2), ms_per_step=round(rm['ms_step'], 4),
                            ns_per_day=round(rm['steps_s'] * 2.0 * 0.0864, 2),
                            speedup_vs_fp64=round(rm['steps_s'] / steps_s, 3),
                            rms_force_err=rms_force_error(sys_, cfg), dtype='f32 pair / f64',
                            near_kernel_ms=round(rm['near_ms'], 4),
                            near_achieved_tflops=round(ach, 3),
                            near_frac_fp32_peak=round(ach / FP32_PEAK_TFLOPS, 4),
                            m2l=m2l_mixed)
        print('bench: mixed %.1f steps/s' % rm['steps_s'], file=sys.stderr, flush=True)
    if not args.no_secondary and name == 'c2':
        rcfg = CONFIGS['ref']
        rsys = make_system('ref')
        sec = dict(workload=rcfg['workload'], boxes=(4 * rcfg['nb2']) ** 3, p=rcfg['p'])
        for prec, key in ((engine.FP64, 'fp64'), (engine.MIXED, 'mixed')):
            if prec == engine.MIXED and args.no_mixed:
                continue
            rr, e = run(rsys, rcfg, prec, args.steps, args.warmup, timer)
            e.close()
            sec[key] = dict(steps_per_s=round(rr['steps_s'], 2), ms_per_step=round(rr['ms_step'], 4),
                            near_kernel_ms=round(rr['near_ms'], 4))
        out['reference_geometry'] = sec
        print('bench: reference geometry done', file=sys.stderr, flush=True)
    if args.drift_steps > 0:
        out['energy_drift'] = dict(fp64=energy_drift(sys_, cfg, engine.FP64, args.drift_steps))
        if not args.no_mixed:
            dm, final = energy_drift(sys_, cfg, engine.MIXED, args.drift_steps, keep=True)
            out['energy_drift']['mixed'] = dm
            out['mixed']['rms_force_err_after_drift'] = rms_force_error(sys_, cfg, final)
        print('bench: drift done', file=sys.stderr, flush=True)
    if args.ifar_steps > 0:
        # the same dynamics with the lists and far field refreshed every 1, 2,
        # 5 and 10 steps, all runs of equal length: how the drift depends on
        # the refresh interval
        out['drift_vs_ifar'] = [energy_drift(sys_, cfg, engine.FP64, args.ifar_steps, ifar=k)
                                for k in (1, 2, 5, 10)]
        print('bench: ifar series done', file=sys.stderr, flush=True)
    if not args.no_ensemble and name == 'c2':
        out['ensemble'] = ensemble_leg(timer, args.ensemble_steps, 10)
        print('bench: ensemble done', file=sys.stderr, flush=True)
    if not args.no_reference and name == 'c2':
        out['reference_workload'] = reference_leg(timer, args.reference_seconds, 200, 20)
        print('bench: reference workload done', file=sys.stderr, flush=True)
    if not args.no_cpu:
        out['cpu_baseline'] = cpu_baseline(sys_, cfg['nb2'], cfg['p'], args.cpu_seconds)
    print(json.dumps(out), flush=True)


if __name__ == '__main__':
    main()
