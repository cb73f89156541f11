"""GPU batched builds of independent molecules: batch = single builds, graph replays =
eager batch on every replay, FP64 parity with the oracle, and the launch-shape
autotuner."""
import numpy as np
import pytest

from oracle import cext as ocx
from oracle import core
from tests import systems

pytestmark = pytest.mark.gpu


def _setup(mols, precision='fp64'):
    import napmo_amd
    grids, bases, plans, Ps = [], [], [], []
    for k, m in enumerate(mols):
        g = napmo_amd.BeckeGrid(napmo_amd.make_species(m['atoms']), None, None, sizes=m['sizes'])
        b = napmo_amd.make_basis(m['atoms'], shells_per_atom=m['shells'])
        f = napmo_amd.GridFock(g, b, m['lmax'], precision)
        S = f.overlap()
        Ps.append(systems.density_matrix(np.random.default_rng(100 + k), S, m['nocc']))
        grids.append(g), bases.append(b), plans.append(f)
    return grids, bases, plans, Ps


def _device(Ps):
    from napmo_amd.device import DeviceArray
    dP = [DeviceArray.from_host(P) for P in Ps]
    dF = [DeviceArray(P.shape) for P in Ps]
    dE = [DeviceArray((2,)) for _ in Ps]
    return dP, dF, dE


def test_batch_matches_single_builds_and_oracle():
    import napmo_amd
    from napmo_amd import workloads
    mols = workloads.batch_molecules(count=6)
    grids, bases, plans, Ps = _setup(mols)
    single = [f.build(P) for f, P in zip(plans, Ps)]
    dP, dF, dE = _device(Ps)
    batch = napmo_amd.BatchFock(plans, nlanes=3)
    batch.build([a.ptr for a in dP], [a.ptr for a in dF], [a.ptr for a in dE])
    batch.synchronize()
    for k, (F, E, Eg) in enumerate(single):
        Fb, Eb = dF[k].to_host(), dE[k].to_host()
        print('molecule %d N=%d n=%d E=%.12f' % (k, grids[k].size, F.shape[0], Eg))
        assert np.array_equal(Fb, F) and Eb[0] == E and Eb[1] == Eg
    # FP64 parity with the oracle on the two smallest molecules
    order = sorted(range(len(mols)), key=lambda k: grids[k].size * mols[k]['nbasis'])
    for k in order[:2]:
        m = mols[k]
        o, ob = systems.oracle_system(m['atoms'], m['shells'], None, None, sizes=m['sizes'])
        phi = ocx.basis_compute(ob, o.points)
        rho = core.density_from_dm(phi, Ps[k])
        v = core.compute_coulomb(o, rho, m['lmax'])
        Fo = core.coupling_matrix(phi, v, o.becke_weights, o.weights)
        Eo = 0.5 * float(np.sum(o.becke_weights * o.weights * rho * v))
        F, E, Eg = single[k]
        print('oracle molecule %d: dE %.3e max|dF| %.3e' % (k, Eg - Eo, np.abs(F - Fo).max()))
        assert abs(Eg - Eo) < 1e-10 * abs(Eo)
        assert np.abs(F - Fo).max() < 1e-10 * np.abs(Fo).max()


def test_batch_graph_replays_match_eager():
    # every replay checked against an eager batch build on the same densities, with eager
    # work and host synchronisation between replays
    import napmo_amd
    from napmo_amd import workloads
    from napmo_amd.system import cext
    mols = workloads.batch_molecules(count=5, seed=6)
    for prec in ('fp64', 'bf16x3'):
        grids, bases, plans, Ps = _setup(mols, prec)
        dP, dF, dE = _device(Ps)
        batch = napmo_amd.BatchFock(plans, nlanes=2)
        ptr = lambda xs: [a.ptr for a in xs]  # noqa: E731
        batch.capture(ptr(dP), ptr(dF), ptr(dE))
        for it in range(3):
            scale = 1.0 + 0.1 * it
            for a, P in zip(dP, Ps):
                Pi = np.ascontiguousarray(P * scale)
                cext.check(cext.napmo_memcpy(a.ptr, Pi.ctypes.data, Pi.nbytes, 1))
            batch.replay()
            batch.synchronize()
            got = [(f.to_host(), e.to_host()) for f, e in zip(dF, dE)]
            batch.build(ptr(dP), ptr(dF), ptr(dE))
            batch.synchronize()
            for k, (Fg, Eg) in enumerate(got):
                assert np.array_equal(Fg, dF[k].to_host()), (prec, it, k)
                assert np.array_equal(Eg, dE[k].to_host()), (prec, it, k)
            print(prec, 'replay', it, 'E', [float(e[1]) for _, e in got[:3]])
        del batch


def test_one_launch_per_phase_matches_single_builds():
    # nlanes = 0: every phase of the batch as one launch (FP64 plans); eager builds and graph
    # replays on changing densities equal the plans' own builds bit for bit; the molecules'
    # lmax values differ, so the shared Poisson tables cut each molecule's harmonics
    import napmo_amd
    from napmo_amd import workloads
    from napmo_amd.system import cext
    mols = workloads.batch_molecules(count=8, seed=12)
    print('lmax per molecule', [m['lmax'] for m in mols])
    assert len({m['lmax'] for m in mols}) > 1
    grids, bases, plans, Ps = _setup(mols)
    dP, dF, dE = _device(Ps)
    batch = napmo_amd.BatchFock(plans, nlanes=0)
    ptr = lambda xs: [a.ptr for a in xs]  # noqa: E731
    batch.build(ptr(dP), ptr(dF), ptr(dE))
    batch.synchronize()
    for k, (f, P) in enumerate(zip(plans, Ps)):
        F, E, Eg = f.build(P)
        Eb = dE[k].to_host()
        assert np.array_equal(dF[k].to_host(), F) and Eb[0] == E and Eb[1] == Eg, k
    batch.capture(ptr(dP), ptr(dF), ptr(dE))
    for it in range(3):
        scale = 1.0 + 0.25 * it
        for a, P in zip(dP, Ps):
            Pi = np.ascontiguousarray(P * scale)
            cext.check(cext.napmo_memcpy(a.ptr, Pi.ctypes.data, Pi.nbytes, 1))
        batch.replay()
        batch.synchronize()
        for k, (f, P) in enumerate(zip(plans, Ps)):
            F, E, Eg = f.build(P * scale)
            Eb = dE[k].to_host()
            assert np.array_equal(dF[k].to_host(), F) and Eb[0] == E and Eb[1] == Eg, (it, k)
        print('replay', it, 'E', [float(e.to_host()[1]) for e in dE[:3]])
    del batch
    # the grouped contraction is FP64 only
    g2, b2, p2, _ = _setup(mols[:2], 'int8')
    with pytest.raises(Exception, match='FP64 plans'):
        napmo_amd.BatchFock(p2, nlanes=0)


def test_autotuner_choice_and_cache(tmp_path):
    import napmo_amd
    from napmo_amd import autotune, workloads
    mols = workloads.batch_molecules(count=1, seed=8)
    for prec in ('fp64', 'bf16x3'):
        grids, bases, plans, Ps = _setup(mols, prec)
        f = plans[0]
        F0, E0, Eg0 = f.build(Ps[0])
        dP, dF, dE = _device(Ps)
        path = str(tmp_path / 'tuning.json')
        setting, src = autotune.tune(f, dP[0].ptr, dF[0].ptr, dE[0].ptr, path=path)
        print(prec, 'tuned', setting, src, autotune.current(f))
        assert src == 'timed'
        F1, E1, Eg1 = f.build(Ps[0])
        assert abs(Eg1 - Eg0) < 1e-12 * abs(Eg0)           # rho and v do not depend on it
        # FP64 kernels differ by summation order only; two BF16x3 kernels (different splits)
        # each stay within 2e-5 max|F| of FP64, so within twice that of each other
        tol = 1e-12 if prec == 'fp64' else 4e-5
        assert np.abs(F1 - F0).max() < tol * np.abs(F0).max()
        setting2, src2 = autotune.tune(f, dP[0].ptr, dF[0].ptr, dE[0].ptr, path=path)
        assert src2 == 'cache' and setting2 == setting


def test_every_launch_shape_gives_the_same_build():
    # all autotuner candidates (Fock workgroup counts, one-pass vs per-tile BF16x3, density
    # stage depth 16 / 32 and 1 / 2 stages in flight) on one molecule: rho and the energy are
    # identical to rounding, F within each precision's tolerance
    from napmo_amd import autotune, workloads
    mols = workloads.batch_molecules(count=1, seed=9)
    for prec in ('fp64', 'bf16x3'):
        grids, bases, plans, Ps = _setup(mols, prec)
        f = plans[0]
        F0, E0, Eg0 = f.build(Ps[0])
        tol = 1e-12 if prec == 'fp64' else 2e-5
        for cand in autotune.candidates(f):
            autotune.apply(f, dict(autotune.DEFAULTS, **cand))
            F1, E1, Eg1 = f.build(Ps[0])
            assert abs(Eg1 - Eg0) < 1e-12 * abs(Eg0), cand
            assert np.abs(F1 - F0).max() < tol * np.abs(F0).max(), cand
        # one or two (point, centre) pairs per lane of the synthesis: the same arithmetic
        # per pair, so the potential and F are bit-identical
        builds = []
        for np_ in (1, 2):
            autotune.apply(f, dict(autotune.DEFAULTS, synth_np=np_))
            builds.append(f.build(Ps[0]))
        assert np.array_equal(builds[0][0], builds[1][0]) and builds[0][2] == builds[1][2]
        print(prec, len(autotune.candidates(f)), 'launch shapes agree')


def test_one_launch_per_phase_single_molecule_and_repeated_pointers():
    # a group of one plan, and a group rebuilt on new output buffers: both equal the plan's
    # own build bit for bit
    import napmo_amd
    from napmo_amd import workloads
    mols = workloads.batch_molecules(count=3, seed=21)
    grids, bases, plans, Ps = _setup(mols)
    for k in range(2):
        dP, dF, dE = _device(Ps[k:k + 1])
        batch = napmo_amd.BatchFock(plans[k:k + 1], nlanes=0)
        batch.build([a.ptr for a in dP], [a.ptr for a in dF], [a.ptr for a in dE])
        batch.synchronize()
        F, E, Eg = plans[k].build(Ps[k])
        Eb = dE[0].to_host()
        assert np.array_equal(dF[0].to_host(), F) and Eb[0] == E and Eb[1] == Eg, k
        dP2, dF2, dE2 = _device(Ps[k:k + 1])
        batch.build([a.ptr for a in dP2], [a.ptr for a in dF2], [a.ptr for a in dE2])
        batch.synchronize()
        assert np.array_equal(dF2[0].to_host(), F), k
        del batch
