#!/usr/bin/env python3
"""Grid Fock build throughput of the nAPMO grid hot path on one MI355X.

One step = one grid Fock build with P resident in HBM: rho = diag(Phi P Phi^T) (MFMA),
Becke-Dickson Poisson (projection, FP64 radial Green solves, synthesis over all
centres), F = Phi^T diag(w bw v) Phi (MFMA) and the FP64 trace E = 1/2 sum P o F.
Phi, the Becke weights and the Poisson plan are per-geometry set-up, outside the step.

Workload (synthetic, seeded): 16 atoms {C,H,N,O} in a 12 bohr box, 50 radial x 590
Lebedev points per atom (N = 472,000), 300 Cartesian s/p/d Gaussians, P = 2 C C^T from
40 S-orthonormal orbitals, lmax = 12.

Prints one JSON line. ``--gpus N`` with N > 1 exits: this build runs on one GPU.
"""
import argparse
import ctypes
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK = {'fp64_mfma': 78.6e12, 'fp64_valu': 78.6e12, 'fp32_mfma': 157.3e12, 'bf16_mfma': 2.5e15,
        'int8_mfma': 5.0e15, 'hbm': 8.0e12}
# Algorithmic work of the synthesis (the measurement plan's 2 M N L): one multiply-add per
# (point, centre, lm) term. The kernel's instruction-level count (Horner spline, 3 FMAs, and
# the harmonic product, 1 FMA: 8 flops per term) is reported beside it.
SYNTH_FLOPS = 2.0
SYNTH_FLOPS_ISSUED = 8.0
# digit products per element of the int8 contraction (pairs i + j <= 6 of 5 digits each)
INT8_PRODUCTS = 15
GATE_EH = 1e-6   # |dE_trace| of a mixed mode against the FP64 mode, in Hartree
METRIC = 'grid Fock builds/s + % roofline, mixed vs FP64 on 1 MI355X at |dE|<=1uEh'


# int8s: the int8 plan with screening: in each chunk of points the 16 x 16 blocks of F bounded
# by 2^-SCREEN_S in every element are skipped (NAPMO_TUNE_INT8_SCREEN), and in the orbital-form
# density each 16-point tile's 16-function groups whose terms of X = Phi C are bounded by
# 2^-SCREEN_S (NAPMO_TUNE_DENSITY_SCREEN)
SCREEN_S = 50
SCREENING = ('F blocks bounded by 2^-%d per element and chunk of points skipped; density terms of '
             'X = Phi C bounded by 2^-%d per 16-point tile and 16-function group skipped')
PLAN_PREC = {'int8s': 'int8'}
MODES = {  # name: (plan, density input)
    'fp64': ('fp64', 'P'),
    'fp64-orb': ('fp64', 'C'),
    'int8-orb': ('int8', 'C'),
    'int8': ('int8', 'P'),
    'int8s-orb': ('int8s', 'C'),
    'bf16x3-orb': ('bf16x3', 'C'),
    'bf16x3': ('bf16x3', 'P'),
}


def gpu_leg(args, log):
    import napmo_amd
    from napmo_amd import workloads
    from napmo_amd.device import DeviceArray
    from napmo_amd.system import cext
    cext.require_gpu()
    atoms, shells, nrad, nang, lmax = workloads.chon16_system()
    t0 = time.perf_counter()
    grid = napmo_amd.BeckeGrid(napmo_amd.make_species(atoms), nrad, nang)
    basis = napmo_amd.make_basis(atoms, shells_per_atom=shells)
    log('grid N=%d basis n=%d set-up %.1fs' % (grid.size, basis['nbasis'], time.perf_counter() - t0))
    modes = ['fp64'] + [m for m in args.modes.split(',') if m and m != 'fp64']
    plans = {}
    for m in modes:
        prec = MODES[m][0]
        if prec not in plans:
            t0 = time.perf_counter()
            try:
                plans[prec] = napmo_amd.GridFock(grid, basis, lmax, PLAN_PREC.get(prec, prec))
                if prec == 'int8s':
                    plans[prec].set_screen(SCREEN_S)
                    plans[prec].set_density_screen(SCREEN_S)
            except cext.NapmoError as e:
                log('precision %s unavailable: %s' % (prec, e))
                continue
            log('plan %s set-up %.2fs' % (prec, time.perf_counter() - t0))
    S = plans['fp64'].overlap()
    C = workloads.occupied_orbitals(np.random.default_rng(2), S, 40)
    P = 2.0 * C @ C.T
    n = basis['nbasis']
    dP = DeviceArray.from_host(P)
    dC = DeviceArray.from_host(C)
    dF = DeviceArray((n, n))
    dE = DeviceArray((2,))
    if not args.no_tune:
        from napmo_amd import autotune
        for prec, plan in plans.items():
            t0 = time.perf_counter()
            if prec == 'fp64':
                # the FP64 plan is tuned on the P-form build of the fp64 mode (the density's
                # stage knobs only act there; fp64-orb shares its Fock contraction shape)
                setting, src = autotune.tune(plan, dP.ptr, dF.ptr, dE.ptr)
            else:
                setting, src = autotune.tune(
                    plan, dP.ptr, dF.ptr, dE.ptr, tag='orb' if prec != 'int8s' else 'orb-s%d' % SCREEN_S,
                    build=lambda plan=plan: plan.build_orbitals_device(dC.ptr, C.shape[1], 2.0, dF.ptr, dE.ptr))
            log('autotune %s: %s (%s, %.1fs)' % (prec, setting, src, time.perf_counter() - t0))
    results = {}
    for m in modes:
        prec, form = MODES[m]
        if prec not in plans:
            continue
        plan = plans[prec]
        if form == 'P':
            step = lambda: plan.build_device(dP.ptr, dF.ptr, dE.ptr)  # noqa: E731
        else:
            step = lambda: plan.build_orbitals_device(dC.ptr, C.shape[1], 2.0, dF.ptr, dE.ptr)  # noqa: E731
        for _ in range(args.warmup):
            step()
        plan.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            step()
        plan.synchronize()
        dt = (time.perf_counter() - t0) / args.steps
        # phase breakdown from a separate pass with HIP events between the phases (the
        # events add a few microseconds per phase, so they stay out of the timed builds)
        plan.profile(args.steps)
        for _ in range(args.steps):
            step()
        plan.synchronize()
        phases, nb = plan.profile_read()
        E = dE.to_host()
        F = dF.to_host()
        results[m] = {'ms_per_build': dt * 1e3, 'builds_per_s': 1.0 / dt, 'E': float(E[0]),
                      'E_grid': float(E[1]), 'phases_ms': phases, 'F': F, 'form': form, 'prec': prec}
        if PLAN_PREC.get(prec, prec) == 'int8':
            kept, total = plan.screen_stats()
            results[m]['blocks_kept'] = kept / total
            log('%s: %d of %d (block, chunk) products computed (%.1f%%)' % (m, kept, total, 100.0 * kept / total))
        log('%s: %.3f ms/build  E_trace=%.12f E_grid=%.12f phases=%s' % (
            m, dt * 1e3, E[0], E[1], {k: round(v, 4) for k, v in phases.items()}))
    N = grid.size
    return results, N, n, lmax, grid.ncenter


def apmo_leg(args, log):
    """Electrons + 4 quantum protons, one grid per species (241,600 + 46,560 points,
    120 + 320 functions): one step = one multi-species Coulomb build (both densities,
    both Poisson expansions, own and cross-species syntheses, both contractions, FP64
    traces), FP64 and BF16x3. Returns one JSON-ready dict."""
    import napmo_amd
    from napmo_amd import workloads
    from napmo_amd.device import DeviceArray
    from napmo_amd.system import cext
    cext.require_gpu()
    sysd = workloads.apmo_system()
    e, nuc = sysd['electron'], sysd['proton']
    grids, bases = [], []
    for sp in (e, nuc):
        grids.append(napmo_amd.BeckeGrid(napmo_amd.make_species(sp['centres']), sp['nrad'], sp['nang']))
        bases.append(napmo_amd.make_basis(sp['centres'], shells_per_atom=sp['shells']))
    rng = np.random.default_rng(3)
    Ps = []
    for sp, g, b in zip((e, nuc), grids, bases):
        S = napmo_amd.GridFock(g, b, sp['lmax'], 'fp64').overlap()
        C = workloads.occupied_orbitals(rng, S, sp['nocc'])
        Ps.append(sp['eta'] * C @ C.T)
    species = [(g, b, sp['charge'], sp['lmax']) for sp, g, b in zip((e, nuc), grids, bases)]
    dP = [DeviceArray.from_host(P) for P in Ps]
    dF = [DeviceArray(P.shape) for P in Ps]
    dE = DeviceArray((8,))
    out = {}
    for prec in ('fp64', 'fp64-one', 'int8', 'int8s', 'bf16x3'):
        base = prec.split('-')[0]
        plan = napmo_amd.APMOFock(species, PLAN_PREC.get(base, base), one_launch=prec.endswith('-one'))
        if prec == 'int8s':
            plan.set_screen(SCREEN_S)
        step = lambda: plan.build_device([a.ptr for a in dP], [a.ptr for a in dF], dE.ptr)  # noqa: E731
        for _ in range(args.warmup):
            step()
        plan.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            step()
        plan.synchronize()
        dt = (time.perf_counter() - t0) / args.steps
        E = dE.to_host().reshape(2, 4)
        out[prec] = {'ms_per_build': dt * 1e3, 'builds_per_s': 1.0 / dt, 'E': E.tolist(),
                     'F': [a.to_host() for a in dF]}
        log('apmo %s: %.3f ms/build  E_e=%s E_n=%s' % (prec, dt * 1e3, E[0], E[1]))
    ref = out['fp64']
    one = out['fp64-one']
    d1 = max(float(np.max(np.abs(a - b))) for a, b in zip(one['F'], ref['F']))
    log('apmo fp64 one launch per phase vs per species: max|dF| %.3e, max dE %.3e' %
        (d1, max(abs(one['E'][s][k] - ref['E'][s][k]) for s in range(2) for k in range(4))))
    best = 'fp64' if ref['builds_per_s'] >= one['builds_per_s'] else 'fp64-one'
    for prec in ('int8', 'int8s', 'bf16x3'):
        mix = out[prec]
        mix['dE_grid'] = max(abs(mix['E'][s][1] - ref['E'][s][1]) for s in range(2))
        mix['dE_trace'] = max(abs(mix['E'][s][0] - ref['E'][s][0]) for s in range(2))
        mix['dF'] = max(float(np.max(np.abs(a - b))) for a, b in zip(mix['F'], ref['F']))
        log('apmo %s vs fp64: max dE_trace %.3e, max dE_grid %.3e, max|dF| %.3e, speed-up %.2fx' %
            (prec, mix['dE_trace'], mix['dE_grid'], mix['dF'], ref['ms_per_build'] / mix['ms_per_build']))
        if mix['dE_trace'] <= GATE_EH and mix['builds_per_s'] > out[best]['builds_per_s']:
            best = prec
    mix = out['int8']
    dEg, dEt, dF = mix['dE_grid'], mix['dE_trace'], mix['dF']
    return {
        'metric': METRIC, 'value': out[best]['builds_per_s'], 'unit': 'builds/s', 'n_gpus': 1,
        'steps': args.steps, 'warmup': args.warmup, 'ms_per_step': out[best]['ms_per_build'],
        'higher_is_better': True, 'scaling': 'weak', 'vs_baseline': None,
        'dtype': {'fp64': 'f64', 'fp64-one': 'f64', 'int8': 'int8+f64', 'int8s': 'int8+f64',
                  'bf16x3': 'bf16x3+f64'}[best],
        'data': 'synthetic (seeded geometry, basis and densities)',
        'config': {'workload': 'multi-species Coulomb build: electrons on 8 x 100 x 302 points (120 '
                               'functions) + 4 quantum protons on 4 x 60 x 194 points (320 functions), '
                               'cross-species Poisson coupling', 'precision': best,
                   'screening': ('F blocks bounded by 2^-%d per element and chunk of points skipped (the '
                                 'P-form density is not screened)' % SCREEN_S if best == 'int8s' else 'none')},
        'modes': {p: {'builds_per_s': r['builds_per_s'], 'ms_per_build': r['ms_per_build'], 'E': r['E'],
                      'dE_trace_vs_fp64': r.get('dE_trace', 0.0), 'dF_max_vs_fp64': r.get('dF', 0.0)}
                  for p, r in out.items()},
        'int8_dE_grid_vs_fp64': dEg, 'int8_dE_trace_vs_fp64': dEt, 'int8_dF_max_vs_fp64': dF,
        'speedup_vs_fp64': ref['ms_per_build'] / out[best]['ms_per_build'],
    }


DPSI_PHASES = ('residual', 'project', 'shell_average', 'radial', 'synth_tables', 'synth',
               'reduce', 'synth_mfma')
SYNTH_MFMA_MIN_KEY = 9          # NAPMO_TUNE_SYNTH_MFMA_MIN


def mfma_densities(grid, nset):
    """How many of nset densities the grid's many-density synthesis puts on MFMA: whole
    tiles of 16 and a last tile of at least the grid's threshold."""
    thr = grid.get_tuning(SYNTH_MFMA_MIN_KEY)
    if thr > 16:
        return 0
    full, rest = 16 * (nset // 16), nset % 16
    return full + (rest if rest >= thr else 0)


def synth_roofline(phases, M, N, L, nset, n_mfma, traffic):
    """Roofline of the dominant synthesis kernel of a many-density workload. MFMA synthesis:
    the product of the pair features (u^c Y_q, 4 L per pair) with the spline coefficients,
    SYNTH_FLOPS_ISSUED M N L flops per density, against the FP64 MFMA peak. One-density
    synthesis: SYNTH_FLOPS M N L per density against the FP64 VALU peak."""
    if n_mfma and phases.get('synth_mfma', 0.0) >= phases.get('synth', 0.0):
        ms = phases['synth_mfma']
        flops = SYNTH_FLOPS_ISSUED * M * N * L * n_mfma
        return {'bound': 'mfma', 'achieved': flops / (ms * 1e-3) / 1e12, 'peak': PEAK['fp64_mfma'] / 1e12,
                'unit': 'TFLOP/s', 'frac': flops / (ms * 1e-3) / PEAK['fp64_mfma'],
                'traffic': table_traffic(traffic, 'k_synth_mfma<'), 'kernel': 'synth_mfma',
                'densities': n_mfma}
    ms = phases['synth']
    flops = SYNTH_FLOPS * M * N * L * (nset - n_mfma)
    return {'bound': 'valu', 'achieved': flops / (ms * 1e-3) / 1e12, 'peak': PEAK['fp64_valu'] / 1e12,
            'unit': 'TFLOP/s', 'frac': flops / (ms * 1e-3) / PEAK['fp64_valu'],
            'traffic': table_traffic(traffic, 'k_synth2n<'), 'kernel': 'synth', 'densities': nset - n_mfma}


def dpsi_orbitals(points, origin, norb, seed=4):
    """Smooth synthetic orbitals on the grid: per orbital a sum of centre Gaussians of
    one exponent with random weights, times a linear factor (no orbital data files)."""
    rng = np.random.default_rng(seed)
    d2 = np.stack([((points - o) ** 2).sum(1) for o in origin])
    psi = np.empty((norb, points.shape[0]))
    for k in range(norb):
        w = rng.uniform(-1.0, 1.0, len(origin))
        a = 0.3 + 0.05 * (k % 7)
        psi[k] = (w[:, None] * np.exp(-a * d2)).sum(0) * (1.0 + 0.1 * (k % 3) * points[:, k % 3])
    return psi


def dpsi_leg(args, log):
    """One Becke-Dickson orbital update step of 40 orbitals on the 16-atom 472,000-point grid
    (lmax 12): residuals R = T + V - e psi, energy corrections, and the corrections dpsi
    (projection, 40 x 16 x 13 radial band solves, synthesis) with psi, T = -1/2 lap psi
    (the GPU kinetic operator) and V = V_nuc psi resident in HBM."""
    import napmo_amd
    from napmo_amd import workloads
    from napmo_amd.device import DeviceArray
    from napmo_amd.system import cext
    from napmo_amd.wavefunction.ndpsi import becke_dickson_step_device
    cext.require_gpu()
    atoms, _, nrad, nang, lmax = workloads.chon16_system()
    norb = 40
    grid = napmo_amd.BeckeGrid(napmo_amd.make_species(atoms), nrad, nang)
    pts, org = grid.points, grid.origin
    N = grid.size
    t0 = time.perf_counter()
    psi = dpsi_orbitals(pts, org, norb)
    T = np.array([napmo_amd.compute_kinetic(grid, p, lmax) for p in psi])
    Z = [workloads.ZNUC[s] for s, _ in atoms]
    Vn = napmo_amd.compute_nuclear(grid, [{'atomic_number': z, 'origin': o} for z, o in zip(Z, org)])
    V = -Vn[None, :] * psi
    e = -0.5 - 0.02 * np.arange(norb)
    log('dpsi inputs: %d orbitals on %d points (%.1fs)' % (norb, N, time.perf_counter() - t0))
    d = {k: DeviceArray.from_host(np.ascontiguousarray(v)) for k, v in
         (('psi', psi), ('T', T), ('V', V), ('e', e))}
    out = {k: DeviceArray(s) for k, s in (('dpsi', (norb, N)), ('de', (norb,)), ('R', (norb, N)))}

    def step():
        becke_dickson_step_device(grid, lmax, norb, d['psi'].ptr, d['T'].ptr, d['V'].ptr, d['e'].ptr,
                                  1.0, out['dpsi'].ptr, out['de'].ptr, out['R'].ptr)

    for _ in range(args.warmup):
        step()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    dt = (time.perf_counter() - t0) / args.steps
    # per-phase HIP-event times in a separate pass (the events add a few us per launch)
    cext.check(cext.napmo_grid_profile(grid._this, 1))
    for _ in range(min(args.steps, 5)):
        step()
    ph = np.zeros(len(DPSI_PHASES))
    nc = ctypes.c_int(0)
    cext.check(cext.napmo_grid_profile_read(grid._this, ph.ctypes.data, len(DPSI_PHASES), ctypes.byref(nc)))
    cext.check(cext.napmo_grid_profile(grid._this, 0))
    phases = dict(zip(DPSI_PHASES, ph.tolist()))
    de = out['de'].to_host()
    dpsi = out['dpsi'].to_host()
    log('dpsi step: %.3f ms (%.1f steps/s), phases %s; de[0:3] %s, max|dpsi| %.3e' % (
        dt * 1e3, 1.0 / dt, {k: round(v, 4) for k, v in phases.items()}, de[:3], np.abs(dpsi).max()))
    M, L = len(atoms), (lmax + 1) ** 2
    # dominant kernel: the synthesis (MFMA tiles and one-density launches)
    name = max(phases, key=phases.get)
    ms = phases[name]
    if name in ('synth', 'synth_mfma'):
        roof = synth_roofline(phases, M, N, L, norb, mfma_densities(grid, norb),
                              traffic_table('traffic_dpsi_chon16.json'))
        name = roof.pop('kernel')
    else:
        nbytes = 8.0 * N * norb * 4
        roof = {'bound': 'hbm', 'achieved': nbytes / (ms * 1e-3) / 1e9, 'peak': PEAK['hbm'] / 1e9,
                'unit': 'GB/s', 'frac': nbytes / (ms * 1e-3) / PEAK['hbm'], 'traffic': None}
    cpu = None
    if not args.no_cpu_baseline:
        cpu = dpsi_cpu_baseline(atoms, nrad, nang, lmax, norb, log)
    return {
        'metric': 'Becke-Dickson orbital update steps/s on 1 MI355X', 'value': 1.0 / dt,
        'unit': 'steps/s', 'n_gpus': 1, 'steps': args.steps, 'warmup': args.warmup,
        'ms_per_step': dt * 1e3, 'higher_is_better': True, 'scaling': 'weak', 'vs_baseline': None,
        'dtype': 'f64', 'data': 'synthetic (seeded orbitals; T from the GPU kinetic operator, V = V_nuc psi)',
        'config': {'workload': 'Becke-Dickson update of 40 orbitals, 16 atoms, 50x590 Becke grid '
                               '(472,000 points), lmax 12', 'norb': norb, 'npoint': N, 'lmax': lmax},
        'roofline': dict(roof, kernel=name), 'cpu_baseline': cpu, 'phases_ms': phases,
    }


def dpsi_cpu_baseline(atoms, nrad, nang, lmax, norb, log):
    """The oracle's compute_dpsi (numpy projection, SciPy sparse radial solves as the
    reference, C/OpenMP synthesis) for one orbital on the full grid, scaled to norb."""
    from oracle import cext as ocx
    from oracle import core
    grid = core.BeckeGrid(atoms, nrad, nang)
    psi = dpsi_orbitals(grid.points, grid.origin, 1)[0]
    r = np.sqrt(((grid.points - grid.origin[0]) ** 2).sum(1))
    V = -psi / np.maximum(r, 1e-3)

    def ev(splines, centre, points):
        out = np.zeros(len(points))
        ocx.eval_decomposition(splines, centre, points, out)
        return out

    t = time.perf_counter()
    core.compute_dpsi(grid, lmax, psi, 0.01, -0.5, 0.1 * psi, V, 1.0, evaluate=ev)
    one = time.perf_counter() - t
    total = one * norb
    log('dpsi cpu baseline (%d threads): %.2fs per orbital, %.1fs per %d-orbital step' % (
        ocx.max_threads(), one, total, norb))
    return {'value': 1.0 / total, 'unit': 'steps/s', 'cores': ocx.max_threads(), 'kind': 'port',
            'sample': 'oracle compute_dpsi of 1 orbital on the full 472,000-point grid (numpy '
                      'projection, SciPy sparse solves, C/OpenMP synthesis), scaled to %d orbitals' % norb,
            'seconds_per_step': total}


def operators_leg(args, log):
    """Grid operators of the numeric-orbital SCF on the 16-atom 472,000-point grid (lmax 12),
    inputs resident in HBM: one exchange-potential build of 10 orbitals (55 Poisson solves
    of orbital products, napmo/wavefunction/psi_numeric.py:186-215) is the step; the
    kinetic operator (nkinetic.py) on the same orbitals is timed beside it."""
    import napmo_amd
    from napmo_amd import workloads
    from napmo_amd.device import DeviceArray
    from napmo_amd.system import cext
    cext.require_gpu()
    atoms, _, nrad, nang, lmax = workloads.chon16_system()
    ndim = 10
    grid = napmo_amd.BeckeGrid(napmo_amd.make_species(atoms), nrad, nang)
    N = grid.size
    psi = dpsi_orbitals(grid.points, grid.origin, ndim)
    D = np.eye(ndim)
    dpsi, dD, dK, dT = (DeviceArray.from_host(np.ascontiguousarray(psi)), DeviceArray.from_host(D),
                        DeviceArray((ndim, N)), DeviceArray((N,)))

    def exchange():
        cext.check(cext.napmo_compute_exchange_potential_device(grid._this, ndim, dpsi.ptr, dD.ptr,
                                                                lmax, dK.ptr))

    def kinetic():
        for k in range(ndim):
            cext.check(cext.napmo_compute_kinetic_device(grid._this, dpsi.ptr + 8 * k * N, lmax, dT.ptr))

    for _ in range(max(1, args.warmup)):
        exchange()
        kinetic()
    steps = max(1, args.steps)
    t0 = time.perf_counter()
    for _ in range(steps):
        exchange()
    dt = (time.perf_counter() - t0) / steps
    t0 = time.perf_counter()
    for _ in range(steps):
        kinetic()
    dt_kin = (time.perf_counter() - t0) / steps / ndim
    cext.check(cext.napmo_grid_profile(grid._this, 1))
    for _ in range(min(steps, 3)):
        exchange()
    ph = np.zeros(len(DPSI_PHASES))
    nc = ctypes.c_int(0)
    cext.check(cext.napmo_grid_profile_read(grid._this, ph.ctypes.data, len(DPSI_PHASES), ctypes.byref(nc)))
    cext.check(cext.napmo_grid_profile(grid._this, 0))
    phases = {'project_radial': float(ph[1]), 'synth_tables': float(ph[4]), 'synth': float(ph[5]),
              'synth_mfma': float(ph[7]), 'reduce_acc': float(ph[6])}
    npair = ndim * (ndim + 1) // 2
    log('exchange build (%d orbitals, %d solves): %.3f ms (%.2f ms per solve); phases %s; '
        'kinetic %.3f ms per orbital' % (ndim, npair, dt * 1e3, dt * 1e3 / npair,
                                         {k: round(v, 3) for k, v in phases.items()}, dt_kin * 1e3))
    M, L = len(atoms), (lmax + 1) ** 2
    roof = synth_roofline(phases, M, N, L, npair, mfma_densities(grid, npair),
                          traffic_table('traffic_dpsi_chon16.json'))
    cpu = None
    if not args.no_cpu_baseline:
        cpu = exchange_cpu_baseline(atoms, nrad, nang, lmax, ndim, log)
    return {
        'metric': 'grid exchange-potential builds/s on 1 MI355X', 'value': 1.0 / dt, 'unit': 'builds/s',
        'n_gpus': 1, 'steps': steps, 'warmup': args.warmup, 'ms_per_step': dt * 1e3,
        'higher_is_better': True, 'scaling': 'weak', 'vs_baseline': None, 'dtype': 'f64',
        'data': 'synthetic (seeded orbitals)',
        'config': {'workload': 'exchange potential of 10 orbitals (55 Poisson solves), 16 atoms, 50x590 '
                               'Becke grid (472,000 points), lmax 12', 'ndim': ndim, 'npoint': N, 'lmax': lmax},
        'roofline': roof, 'cpu_baseline': cpu, 'phases_ms': phases, 'kinetic_ms_per_orbital': dt_kin * 1e3,
    }


def exchange_cpu_baseline(atoms, nrad, nang, lmax, ndim, log):
    """The oracle's compute_coulomb of one orbital product on the full grid (numpy
    projection, SciPy sparse solves, C/OpenMP synthesis), scaled to the ndim(ndim+1)/2
    solves of an exchange build."""
    from oracle import cext as ocx
    from oracle import core
    grid = core.BeckeGrid(atoms, nrad, nang)
    psi = dpsi_orbitals(grid.points, grid.origin, 2)
    t = time.perf_counter()
    U = core.poisson_solver(grid, psi[0] * psi[1], lmax)
    out = np.zeros(grid.size)
    for k in range(grid.ncenter):
        ocx.eval_decomposition(U[k], grid.origin[k], grid.points, out)
    one = time.perf_counter() - t
    total = one * ndim * (ndim + 1) // 2
    log('exchange cpu baseline (%d threads): %.2fs per solve, %.1fs per build' % (ocx.max_threads(), one, total))
    return {'value': 1.0 / total, 'unit': 'builds/s', 'cores': ocx.max_threads(), 'kind': 'port',
            'sample': 'oracle Poisson solve of 1 orbital product on the full 472,000-point grid (numpy '
                      'projection, SciPy sparse solves, C/OpenMP synthesis), scaled to %d solves'
                      % (ndim * (ndim + 1) // 2), 'seconds_per_step': total}


def batch_leg(args, log):
    """64 independent molecules (2-12 atoms, uneven 35-110 x 110-590 grids, 20-200
    functions): one step = one grid Fock build of every molecule, launched as one batch
    over --lanes streams, or (fp64-group) with one launch per phase for all molecules, and
    replayed from a captured HIP graph. FP64 and int8, from the density matrices (P form) or
    the occupied orbitals (-orb, P = 2 C C^T), launch shapes from the autotuner; per-molecule
    |dE| of every other mode against the FP64 P-form build."""
    import napmo_amd
    from napmo_amd import autotune, workloads
    from napmo_amd.device import DeviceArray
    from napmo_amd.system import cext
    cext.require_gpu()
    mols = workloads.batch_molecules()
    t0 = time.perf_counter()
    grids, bases, Ps = [], [], []
    for k, m in enumerate(mols):
        g = napmo_amd.BeckeGrid(napmo_amd.make_species(m['atoms']), None, None, sizes=m['sizes'])
        b = napmo_amd.make_basis(m['atoms'], shells_per_atom=m['shells'])
        grids.append(g)
        bases.append(b)
    log('batch: %d molecules, %d points, grids in %.1fs' % (len(mols), sum(g.size for g in grids),
                                                           time.perf_counter() - t0))
    dP, dC, dF, dE, Cs = [], [], [], [], []
    out = {}
    for mode in ('fp64', 'fp64-group', 'int8', 'fp64-orb', 'int8-orb', 'int8s-orb'):
        prec = mode.split('-')[0]
        lanes = 0 if mode.endswith('-group') else args.lanes
        orb = mode.endswith('-orb')
        t0 = time.perf_counter()
        plans = [napmo_amd.GridFock(g, b, m['lmax'], PLAN_PREC.get(prec, prec)) for g, b, m in zip(grids, bases, mols)]
        if prec == 'int8s':
            for f in plans:
                f.set_screen(SCREEN_S)
                f.set_density_screen(SCREEN_S)
        if not Ps:
            rng = np.random.default_rng(5)
            for m, f in zip(mols, plans):
                C = workloads.occupied_orbitals(rng, f.overlap(), m['nocc'])
                Cs.append(C)
                Ps.append(2.0 * C @ C.T)
            dP = [DeviceArray.from_host(P) for P in Ps]
            dC = [DeviceArray.from_host(C) for C in Cs]
            dF = [DeviceArray(P.shape) for P in Ps]
            dE = [DeviceArray((2,)) for _ in Ps]
        nocc = [C.shape[1] for C in Cs]
        sources = {}
        for f, a, c_, b_, e_, k in zip(plans, dP, dC, dF, dE, nocc):
            if orb:
                _, src = autotune.tune(f, a.ptr, b_.ptr, e_.ptr,
                                       tag='orb' if prec != 'int8s' else 'orb-s%d' % SCREEN_S,
                                       build=lambda f=f, c_=c_, b_=b_, e_=e_, k=k:
                                       f.build_orbitals_device(c_.ptr, k, 2.0, b_.ptr, e_.ptr))
            else:
                _, src = autotune.tune(f, a.ptr, b_.ptr, e_.ptr)
            sources[src] = sources.get(src, 0) + 1
        log('batch %s: plans + tuning %.1fs (%s)' % (mode, time.perf_counter() - t0, sources))
        t0 = time.perf_counter()
        batch = napmo_amd.BatchFock(plans, nlanes=lanes)
        if lanes == 0:
            log('batch %s: one-launch-per-phase set-up %.1fs' % (mode, time.perf_counter() - t0))
        ptr = lambda xs: [x.ptr for x in xs]  # noqa: E731
        if orb:
            step = lambda: batch.build_orbitals(ptr(dC), nocc, 2.0, ptr(dF), ptr(dE))  # noqa: E731
            record = lambda: batch.capture_orbitals(ptr(dC), nocc, 2.0, ptr(dF), ptr(dE))  # noqa: E731
        else:
            step = lambda: batch.build(ptr(dP), ptr(dF), ptr(dE))  # noqa: E731
            record = lambda: batch.capture(ptr(dP), ptr(dF), ptr(dE))  # noqa: E731
        for _ in range(args.warmup):
            step()
        batch.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            step()
        batch.synchronize()
        t_eager = (time.perf_counter() - t0) / args.steps
        record()
        for _ in range(args.warmup):
            batch.replay()
        batch.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            batch.replay()
        batch.synchronize()
        t_graph = (time.perf_counter() - t0) / args.steps
        E = np.array([e.to_host() for e in dE])
        out[mode] = {'ms_eager': t_eager * 1e3, 'ms_graph': t_graph * 1e3, 'E': E,
                     'F': [f.to_host() for f in dF]}
        log('batch %s: %.3f ms per batch eager, %.3f ms graph replay (%.1f molecule builds/s)' %
            (mode, t_eager * 1e3, t_graph * 1e3, len(mols) / t_graph))
        del batch, plans
    same = (np.array_equal(out['fp64-group']['E'], out['fp64']['E']) and
            all(np.array_equal(a, b) for a, b in zip(out['fp64-group']['F'], out['fp64']['F'])))
    log('batch fp64-group = fp64 over streams, bit for bit: %s' % same)
    if not same:
        raise RuntimeError('one-launch-per-phase batch differs from the per-plan builds')
    ref = out['fp64']
    gate = {}
    for mode in ('int8', 'fp64-orb', 'int8-orb', 'int8s-orb'):
        r = out[mode]
        r['dEg'] = np.abs(r['E'][:, 1] - ref['E'][:, 1])
        r['dEt'] = np.abs(r['E'][:, 0] - ref['E'][:, 0])
        r['dF'] = [float(np.max(np.abs(a - b))) for a, b in zip(r['F'], ref['F'])]
        gate[mode] = r['dEt'].max() <= GATE_EH
        log('batch %s vs fp64 per molecule: max |dE_trace| %.3e, max |dE_grid| %.3e, max |dF| %.3e' %
            (mode, r['dEt'].max(), r['dEg'].max(), max(r['dF'])))
    ok = ['fp64', 'fp64-group'] + [m for m, g in gate.items() if g]
    best = min(ok, key=lambda k: out[k]['ms_graph'])
    t = out[best]['ms_graph'] * 1e-3
    return {
        'metric': METRIC, 'value': len(mols) / t, 'unit': 'molecule builds/s', 'n_gpus': 1,
        'steps': args.steps, 'warmup': args.warmup, 'ms_per_step': t * 1e3, 'higher_is_better': True,
        'scaling': 'weak', 'vs_baseline': None, 'dtype': 'int8+f64' if best.startswith('int8') else 'f64',
        'data': 'synthetic (seeded molecules, bases and orbitals)',
        'config': {'workload': 'batch of 64 molecules, 2-12 atoms, uneven 35-110 x 110-590 grids '
                               '(%d points), 20-200 functions, HIP-graph replay (%s)'
                               % (sum(g.size for g in grids),
                                  'one launch per phase' if best.endswith('-group')
                                  else '%d streams' % args.lanes), 'precision': best,
                   'screening': (SCREENING % (SCREEN_S, SCREEN_S) if best.startswith('int8s') else 'none')},
        'modes': {p: {'ms_per_batch_eager': r['ms_eager'], 'ms_per_batch_graph': r['ms_graph']}
                  for p, r in out.items()},
        'per_molecule_vs_fp64': {m: {'dE_grid': out[m]['dEg'].tolist(), 'dE_trace': out[m]['dEt'].tolist(),
                                     'dF_max': out[m]['dF']} for m in gate},
        'speedup_vs_fp64': ref['ms_graph'] * 1e-3 / t,
    }


PHASE_KERNELS = {'density': 'k_density_mfma',
                 'fock': {'bf16x3': 'k_fock_bf16x3_pass', 'fp64': 'k_fock_f64', 'int8': 'k_ozaki_fock',
                          'int8s': 'k_ozaki_fock'},
                 'synth': 'k_synth2n<', 'project': 'k_project', 'radial': 'k_radial',
                 'fock_reduce': 'k_fock_reduce'}


def traffic_table(name):
    """Per-kernel HBM bytes of a committed rocprofv3 FETCH_SIZE / WRITE_SIZE summary."""
    path = os.path.join(ROOT, 'profiles', name)
    if not os.path.exists(path):
        return {}
    with open(path) as f:
        return json.load(f).get('kernels', {})


def table_traffic(kern, want):
    """Read + write bytes per launch of the first kernel whose name contains want."""
    for k in kern:
        if want in k:
            r, w = kern[k].get('read_bytes'), kern[k].get('write_bytes')
            if r is not None and w is not None:
                return r + w
    return None


def measured_traffic(res, phase):
    """HBM bytes per launch (read + write) of the phase's kernel from the committed
    rocprofv3 FETCH_SIZE / WRITE_SIZE summary (profiles/traffic_chon16.json; the screened int8
    mode's contraction from profiles/traffic_chon16_int8_screened.json), or None."""
    kern = traffic_table('traffic_chon16_int8_screened.json' if res['prec'] == 'int8s' and phase == 'fock'
                         else 'traffic_chon16.json')
    if not kern:
        return None
    want = PHASE_KERNELS.get(phase)
    if want is None:
        return None
    if isinstance(want, dict):
        want = want[res['prec']]
    names = [k for k in kern if want in k]
    if phase == 'density':
        # orbital form: the streamed kernel (C resident in LDS) when the profile has it
        stream = [k for k in kern if 'k_density_orb_stream' in k]
        tag = ', 1, ' if res['form'] == 'C' else ', 0, '
        names = (stream if res['form'] == 'C' and stream else
                 [k for k in names if tag in k] or names)
    for k in names:
        r, w = kern[k].get('read_bytes'), kern[k].get('write_bytes')
        if r is not None and w is not None:
            return r + w
    return None


def phase_work(res, name, N, n, L, M):
    """(work, peak, unit, bound) of a phase per the measurement plan: density 2 N n nocc (orbital
    form) or 2 N n^2 on FP64 MFMA; F 2 N n^2 (FP64 MFMA; int8: INT8_PRODUCTS digit products of
    it on int8 MFMA; bf16x3: 3 bf16 products); projection 2 N L and synthesis 2 M N L on the FP64
    vector units; other phases by their HBM bytes."""
    if name == 'density':
        k = n if res['form'] == 'P' else res['nocc']
        return 2.0 * N * n * k, PEAK['fp64_mfma'], 'TFLOP/s', 'mfma'
    if name == 'fock':
        if res['prec'] in ('int8', 'int8s'):
            # screened: the (block, chunk) products the kernel computed
            return (INT8_PRODUCTS * 2.0 * N * n * n * res.get('blocks_kept', 1.0), PEAK['int8_mfma'],
                    'TOP/s', 'mfma')
        if res['prec'] == 'bf16x3':
            return 3 * 2.0 * N * n * n, PEAK['bf16_mfma'], 'TFLOP/s', 'mfma'
        return 2.0 * N * n * n, PEAK['fp64_mfma'], 'TFLOP/s', 'mfma'
    if name == 'synth':
        return SYNTH_FLOPS * M * N * L, PEAK['fp64_valu'], 'TFLOP/s', 'valu'
    if name == 'project':
        return 2.0 * N * L, PEAK['fp64_valu'], 'TFLOP/s', 'valu'
    return 8.0 * N * 4, PEAK['hbm'], 'GB/s', 'hbm'


def roofline(res, N, n, L, M):
    """Roofline of the dominant phase: algorithmic work per launch / its mean HIP-event time."""
    ph = res['phases_ms']
    name = max(ph, key=ph.get)
    ms = ph[name]
    work, peak, unit, bound = phase_work(res, name, N, n, L, M)
    scale = 1e9 if unit == 'GB/s' else 1e12
    roof = {'bound': bound, 'achieved': work / (ms * 1e-3) / scale, 'peak': peak / scale,
            'unit': unit, 'frac': work / (ms * 1e-3) / peak, 'traffic': measured_traffic(res, name)}
    if name == 'fock' and res['prec'] == 'int8s':
        roof['blocks_kept'] = res['blocks_kept']
        roof['achieved_dense_equivalent'] = roof['achieved'] / res['blocks_kept']
    if name == 'synth':
        roof['issued_flops_per_term'] = SYNTH_FLOPS_ISSUED
        roof['achieved_issued'] = SYNTH_FLOPS_ISSUED * M * N * L / (ms * 1e-3) / 1e12
    return name, roof


def build_roofline(res, N, n, L, M):
    """Whole-build roofline time per the measurement plan: sum of the compute phases at their
    peaks (density and F on MFMA, projection and synthesis on the FP64 vector units), or the
    algorithmic HBM bytes (Phi read twice at its stored width, the grid arrays, P and F), if
    larger; fraction = t_roof / t_measured."""
    t = 0.0
    for name in ('density', 'fock', 'project', 'synth'):
        work, peak, _, _ = phase_work(res, name, N, n, L, M)
        t += work / peak
    phi_bytes = {'fp64': 8.0, 'int8': 4.0, 'int8s': 4.0, 'bf16x3': 4.0}[res['prec']]
    nbytes = N * n * 8.0 + N * n * phi_bytes + 40.0 * N + 16.0 * n * n
    t = max(t, nbytes / PEAK['hbm'])
    return {'t_roof_ms': t * 1e3, 't_ms': res['ms_per_build'], 'frac': t * 1e3 / res['ms_per_build']}


def cpu_baseline(budget_s, log):
    """The oracle (C/OpenMP restatement of the reference hot path, Python orchestration
    of the Poisson solves as napmo/grids/poisson_solver.py) timed per stage on a bounded
    sample of the 16-atom, 472,000-point build and scaled to one build, with all host
    threads and with one thread."""
    from oracle import cext as ocx
    from oracle import core
    from napmo_amd import workloads
    atoms, shells, nrad, nang, lmax = workloads.chon16_system()
    threads = ocx.max_threads()
    t0 = time.perf_counter()
    grid = core.BeckeGrid(atoms, nrad, nang)
    basis = core.Basis()
    for (s, o), sh in zip(atoms, shells):
        for shell in sh:
            basis.add_shell(o, shell['l'], shell['prim'], shell['cont'])
    N, n, M = grid.size, basis.nbasis, grid.ncenter
    rng = np.random.default_rng(0)
    P = rng.normal(size=(n, n))
    P = P + P.T
    # Poisson for one atom (projection + all (l, m) solves, single-threaded Python and
    # SciPy as in the reference), scaled by the atom count
    ag = grid.atgrids[0]
    f = np.abs(rng.normal(size=ag.size))
    t = time.perf_counter()
    sph = core.spherical_expansion(ag, f, lmax)
    U = core.poisson_radial(ag, sph, lmax)
    t_poisson = (time.perf_counter() - t) * M

    def stages(ns, nthreads):
        ocx.set_threads(nthreads)
        idx = np.sort(rng.choice(N, ns, replace=False))
        pts = grid.points[idx]
        phi = ocx.basis_compute(basis, pts)
        t = time.perf_counter()
        ocx.density(phi, P)
        t_density = (time.perf_counter() - t) * N / ns
        t = time.perf_counter()
        out = np.zeros(ns)
        ocx.eval_decomposition(U, grid.origin[0], pts, out)
        t_synth = (time.perf_counter() - t) * M * N / ns
        v = rng.normal(size=ns)
        t = time.perf_counter()
        ocx.coupling(np.ascontiguousarray(phi.T), v, grid.becke_weights[idx], grid.weights[idx])
        t_fock = (time.perf_counter() - t) * N / ns
        return t_density, t_synth, t_fock

    # warm-up, then the median of 5 samples of every stage (all host threads), and one
    # single-thread sample
    stages(4096, threads)
    reps = [stages(16384, threads) for _ in range(5)]
    d_all, s_all, f_all = (float(np.median([r[k] for r in reps])) for k in range(3))
    pois = []
    for _ in range(3):
        t = time.perf_counter()
        core.poisson_radial(ag, core.spherical_expansion(ag, f, lmax), lmax)
        pois.append((time.perf_counter() - t) * M)
    t_poisson = float(np.median(pois + [t_poisson]))
    d_one, s_one, f_one = stages(2048, 1)
    ocx.set_threads(threads)
    total = d_all + t_poisson + s_all + f_all
    total_one = d_one + t_poisson + s_one + f_one
    model = ''
    try:
        with open('/proc/cpuinfo') as fh:
            model = next((l.split(':', 1)[1].strip() for l in fh if l.startswith('model name')), '')
    except OSError:
        pass
    log('cpu baseline (%d threads, %s): density %.2fs poisson %.2fs synth %.2fs fock %.2fs per build '
        '(median of 5); 1 thread: %.1fs per build (sampled in %.1fs)' % (
            threads, model, d_all, t_poisson, s_all, f_all, total_one, time.perf_counter() - t0))
    return {'value': 1.0 / total, 'unit': 'builds/s', 'cores': threads, 'kind': 'port', 'cpu_model': model,
            'sample': 'oracle C/OpenMP stages on 16384 of 472000 points (density, synthesis over 1 of '
                      '16 centres, contraction; median of 5 after a warm-up) and the full Poisson of 1 of '
                      '16 atoms (median of 4), each scaled to one 16-atom 472,000-point build; '
                      'Becke/collocation set-up excluded as on the GPU',
            'seconds_per_build': total,
            'one_thread': {'value': 1.0 / total_one, 'seconds_per_build': total_one,
                           'sample': 'same stages on 2048 points, 1 thread'}}


def main():
    ap = argparse.ArgumentParser(description=__doc__.split('\n')[0])
    ap.add_argument('--gpus', type=int, default=1)
    ap.add_argument('--steps', type=int, default=20)
    ap.add_argument('--warmup', type=int, default=3)
    ap.add_argument('--modes', default='fp64,fp64-orb,int8-orb,int8s-orb,bf16x3-orb',
                    help='comma list of modes (%s); fp64 is always timed' % ', '.join(MODES))
    ap.add_argument('--no-cpu-baseline', action='store_true')
    ap.add_argument('--no-tune', action='store_true', help='keep the default launch shapes')
    ap.add_argument('--lanes', type=int, default=4, help='HIP streams of the batch64 workload (the device has 4 hardware queues)')
    ap.add_argument('--copy-calibration', type=float, default=0.0, metavar='GIB',
                    help='only run a plain device copy of GIB gibibytes 5 times (counter calibration)')
    ap.add_argument('--cpu-seconds', type=float, default=30.0)
    ap.add_argument('--workload', default='chon16',
                    choices=('chon16', 'apmo', 'batch64', 'dpsi', 'operators'),
                    help='chon16: 16-atom 472,000-point build (default); apmo: electrons + '
                         'quantum protons on separate grids; batch64: 64 small molecules; dpsi: '
                         'Becke-Dickson update of 40 orbitals on the 16-atom grid; operators: '
                         'exchange potential of 10 orbitals and the kinetic operator')
    args = ap.parse_args()

    def log(msg):
        print('[bench] ' + msg, file=sys.stderr, flush=True)

    if args.copy_calibration > 0:
        from napmo_amd.device import DeviceArray
        from napmo_amd.system import cext
        n = int(args.copy_calibration * (1 << 30) / 8) // 2 * 2
        src, dst = DeviceArray((n,)), DeviceArray((n,))
        for _ in range(5):
            cext.check(cext.napmo_device_copy(dst.ptr, src.ptr, n))
        print(json.dumps({'copy_bytes_read': n * 8, 'copy_bytes_written': n * 8, 'launches': 5}))
        return
    if args.gpus != 1:
        print('bench.py: this build runs the grid hot path on one GPU; --gpus %d is not supported '
              '(the sharding design is documented, not built)' % args.gpus, file=sys.stderr)
        sys.exit(2)
    if args.workload == 'apmo':
        print(json.dumps(apmo_leg(args, log)), flush=True)
        return
    if args.workload == 'batch64':
        print(json.dumps(batch_leg(args, log)), flush=True)
        return
    if args.workload == 'dpsi':
        print(json.dumps(dpsi_leg(args, log)), flush=True)
        return
    if args.workload == 'operators':
        print(json.dumps(operators_leg(args, log)), flush=True)
        return
    results, N, n, lmax, M = gpu_leg(args, log)
    L = (lmax + 1) ** 2
    ref = results['fp64']
    best = 'fp64'
    for m, r in results.items():
        r['nocc'] = 40
        if m == 'fp64':
            continue
        # the energy of the F each mode returns (1/2 tr(P F)) against the FP64 build's; the FP64
        # grid trace 1/2 sum w bw rho v is reported beside it
        r['dE_trace'] = r['E'] - ref['E']
        r['dE_grid'] = r['E_grid'] - ref['E_grid']
        r['dF_max'] = float(np.max(np.abs(r['F'] - ref['F'])))
        log('%s vs fp64: dE_trace = %.3e Eh, dE_grid = %.3e Eh, max|dF| = %.3e Eh, speed-up %.2fx' %
            (m, r['dE_trace'], r['dE_grid'], r['dF_max'], ref['ms_per_build'] / r['ms_per_build']))
        if abs(r['dE_trace']) <= GATE_EH and r['builds_per_s'] > results[best]['builds_per_s']:
            best = m
    head = results[best]
    kname, roof = roofline(head, N, n, L, M)
    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(args.cpu_seconds, log)
    fp64_same_form = 'fp64-orb' if head['form'] == 'C' and 'fp64-orb' in results else 'fp64'
    line = {
        'metric': METRIC,
        'value': head['builds_per_s'],
        'unit': 'builds/s',
        'n_gpus': 1,
        'steps': args.steps,
        'warmup': args.warmup,
        'ms_per_step': head['ms_per_build'],
        'higher_is_better': True,
        'scaling': 'weak',
        'vs_baseline': None,
        'dtype': {'fp64': 'f64', 'int8': 'int8+f64', 'int8s': 'int8+f64', 'bf16x3': 'bf16x3+f64'}[head['prec']],
        'data': 'synthetic (seeded geometry, basis and density; no reference data files)',
        'config': {'workload': 'grid Fock build, 16 atoms, 50x590 Becke grid (472,000 points), '
                               '300 s/p/d Gaussians, lmax 12, P resident in HBM',
                   'npoint': N, 'nbasis': n, 'ncenter': M, 'lmax': lmax, 'precision': best,
                   'screening': (SCREENING % (SCREEN_S, SCREEN_S) if head['prec'] == 'int8s' else 'none')},
        'roofline': dict(roof, kernel=kname),
        'build_roofline': build_roofline(head, N, n, L, M),
        'cpu_baseline': cpu,
        'gate_dE_trace_eh': GATE_EH,
        'dE_trace_vs_fp64': head.get('dE_trace', 0.0),
        'dF_max_vs_fp64': head.get('dF_max', 0.0),
        'modes': {m: {'builds_per_s': r['builds_per_s'], 'ms_per_build': r['ms_per_build'],
                      'E_grid': r['E_grid'], 'E_trace': r['E'], 'dE_grid_vs_fp64': r.get('dE_grid', 0.0),
                      'dE_trace_vs_fp64': r.get('dE_trace', 0.0), 'dF_max_vs_fp64': r.get('dF_max', 0.0),
                      'meets_gate': abs(r.get('dE_trace', 0.0)) <= GATE_EH,
                      'blocks_kept': r.get('blocks_kept'),
                      'phases_ms': r['phases_ms'],
                      'build_roofline': build_roofline(r, N, n, L, M)}
                  for m, r in results.items()},
        'speedup_vs_fp64': ref['ms_per_build'] / head['ms_per_build'],
        'speedup_vs_fp64_same_density_form': results[fp64_same_form]['ms_per_build'] / head['ms_per_build'],
    }
    print(json.dumps(line), flush=True)


if __name__ == '__main__':
    main()
