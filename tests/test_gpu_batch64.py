"""The 64-molecule batch at its full size (64 seeded molecules, 2-12 atoms, uneven 35-110 x
110-590 grids, 20-200 functions): batch = single builds bit for bit, HIP-graph replays =
eager batch, FP64 parity with the oracle on the smallest molecules, and the int8 batch's
per-molecule energy error against the FP64 build within 1e-6 Eh (also with the int8
contraction's block screening at 2^-50)."""
import numpy as np
import pytest

from oracle import cext as ocx
from oracle import core
from tests import systems

pytestmark = pytest.mark.gpu

GATE_EH = 1e-6


@pytest.fixture(scope='module')
def batch64():
    import napmo_amd
    from napmo_amd import workloads
    mols = workloads.batch_molecules()
    assert len(mols) == 64
    grids = [napmo_amd.BeckeGrid(napmo_amd.make_species(m['atoms']), None, None, sizes=m['sizes'])
             for m in mols]
    bases = [napmo_amd.make_basis(m['atoms'], shells_per_atom=m['shells']) for m in mols]
    rng = np.random.default_rng(5)
    Ps = []
    for m, g, b in zip(mols, grids, bases):
        f = napmo_amd.GridFock(g, b, m['lmax'], 'fp64')
        C = workloads.occupied_orbitals(rng, f.overlap(), m['nocc'])
        Ps.append(2.0 * C @ C.T)
        del f
    print('batch64: %d points, n = %d..%d' % (sum(g.size for g in grids),
                                             min(m['nbasis'] for m in mols), max(m['nbasis'] for m in mols)))
    return mols, grids, bases, Ps


def _device(Ps):
    from napmo_amd.device import DeviceArray
    return ([DeviceArray.from_host(P) for P in Ps], [DeviceArray(P.shape) for P in Ps],
            [DeviceArray((2,)) for _ in Ps])


def _ptr(xs):
    return [x.ptr for x in xs]


def _run(batch64, prec):
    """Single builds, an eager batch over 4 streams and three graph replays (eager work and
    host synchronisation between them), each replay checked against the eager batch."""
    import napmo_amd
    from napmo_amd.system import cext
    mols, grids, bases, Ps = batch64
    plans = [napmo_amd.GridFock(g, b, m['lmax'], prec) for g, b, m in zip(grids, bases, mols)]
    single = [f.build(P) for f, P in zip(plans, Ps)]
    dP, dF, dE = _device(Ps)
    batch = napmo_amd.BatchFock(plans, nlanes=4)
    batch.build(_ptr(dP), _ptr(dF), _ptr(dE))
    batch.synchronize()
    for k, (F, E, Eg) in enumerate(single):
        Fb, Eb = dF[k].to_host(), dE[k].to_host()
        assert np.array_equal(Fb, F) and Eb[0] == E and Eb[1] == Eg, (prec, k)
    eager = [(f.to_host(), e.to_host()) for f, e in zip(dF, dE)]
    batch.capture(_ptr(dP), _ptr(dF), _ptr(dE))
    for it in range(3):
        for a, P in zip(dP, Ps):  # eager copies of the same densities between replays
            Pi = np.ascontiguousarray(P)
            cext.check(cext.napmo_memcpy(a.ptr, Pi.ctypes.data, Pi.nbytes, 1))
        for a, P in zip(dF, Ps):  # outputs cleared, so every replay rewrites them
            Z = np.zeros_like(P)
            cext.check(cext.napmo_memcpy(a.ptr, Z.ctypes.data, Z.nbytes, 1))
        batch.replay()
        batch.synchronize()
        for k, (Fe, Ee) in enumerate(eager):
            assert np.array_equal(dF[k].to_host(), Fe), (prec, it, k)
            assert np.array_equal(dE[k].to_host(), Ee), (prec, it, k)
    print('%s: 64 single builds = eager batch = 3 graph replays' % prec)
    del batch, plans
    return single


@pytest.fixture(scope='module')
def fp64_builds(batch64):
    return _run(batch64, 'fp64')


def test_batch64_fp64_batch_graph_and_oracle(batch64, fp64_builds):
    mols, grids, bases, Ps = batch64
    single = fp64_builds
    # FP64 parity with the oracle on the three smallest molecules (points x functions)
    order = sorted(range(len(mols)), key=lambda k: grids[k].size * mols[k]['nbasis'])
    for k in order[:3]:
        m = mols[k]
        o, ob = systems.oracle_system(m['atoms'], m['shells'], None, None, sizes=m['sizes'])
        phi = ocx.basis_compute(ob, o.points)
        rho = core.density_from_dm(phi, Ps[k])
        v = core.compute_coulomb(o, rho, m['lmax'])
        Fo = core.coupling_matrix(phi, v, o.becke_weights, o.weights)
        Eo = 0.5 * float(np.sum(o.becke_weights * o.weights * rho * v))
        F, E, Eg = single[k]
        print('oracle molecule %d (N=%d, n=%d): dE %.3e max|dF| %.3e' %
              (k, grids[k].size, m['nbasis'], Eg - Eo, np.abs(F - Fo).max()))
        assert abs(Eg - Eo) < 1e-10 * abs(Eo)
        assert abs(E - Eo) < 1e-10 * abs(Eo)
        assert np.abs(F - Fo).max() < 1e-10 * np.abs(Fo).max()


def test_batch64_int8_per_molecule_energy(batch64, fp64_builds):
    # the int8 batch (exact int32 accumulation of 28-bit digit slices, FP64 chunk sums):
    # batch = single builds = graph replays, and every molecule's 1/2 tr(P F) within 1e-6 Eh
    # of the FP64 build's
    mols, grids, bases, Ps = batch64
    mixed = _run(batch64, 'int8')
    dEt = np.array([abs(m[1] - r[1]) for m, r in zip(mixed, fp64_builds)])
    dEg = np.array([abs(m[2] - r[2]) for m, r in zip(mixed, fp64_builds)])
    dF = np.array([np.abs(m[0] - r[0]).max() / np.abs(r[0]).max() for m, r in zip(mixed, fp64_builds)])
    print('int8 batch64: max |dE_trace| %.3e Eh (median %.3e), max |dE_grid| %.3e, max rel |dF| %.3e'
          % (dEt.max(), np.median(dEt), dEg.max(), dF.max()))
    assert dEg.max() <= 1e-12 * max(abs(r[2]) for r in fp64_builds)  # rho, v FP64 in every mode
    assert dEt.max() <= GATE_EH
    assert dF.max() < 1e-7


def test_batch64_one_launch_per_phase(batch64, fp64_builds):
    # nlanes = 0: the 64 FP64 builds with one launch per phase (the molecules' atoms and
    # points side by side, each molecule keeping its own lmax 8..12 and its own pairs):
    # equal to the single builds bit for bit, eagerly and from three graph replays
    import napmo_amd
    from napmo_amd.system import cext
    mols, grids, bases, Ps = batch64
    plans = [napmo_amd.GridFock(g, b, m['lmax'], 'fp64') for g, b, m in zip(grids, bases, mols)]
    dP, dF, dE = _device(Ps)
    batch = napmo_amd.BatchFock(plans, nlanes=0)
    batch.build(_ptr(dP), _ptr(dF), _ptr(dE))
    batch.synchronize()
    for k, (F, E, Eg) in enumerate(fp64_builds):
        Fb, Eb = dF[k].to_host(), dE[k].to_host()
        assert np.array_equal(Fb, F) and Eb[0] == E and Eb[1] == Eg, k
    batch.capture(_ptr(dP), _ptr(dF), _ptr(dE))
    for it in range(3):
        for a, P in zip(dF, Ps):  # outputs cleared, so every replay rewrites them
            Z = np.zeros_like(P)
            cext.check(cext.napmo_memcpy(a.ptr, Z.ctypes.data, Z.nbytes, 1))
        batch.replay()
        batch.synchronize()
        for k, (F, E, Eg) in enumerate(fp64_builds):
            Eb = dE[k].to_host()
            assert np.array_equal(dF[k].to_host(), F) and Eb[0] == E and Eb[1] == Eg, (it, k)
    print('fp64, one launch per phase: 64 builds = single builds, eager and 3 graph replays')
    del batch, plans


def test_batch64_orbital_form(batch64, fp64_builds):
    # the orbital form over 4 streams (P = 2 C C^T from the same seeded orbitals): the batch
    # equals the plans' own orbital-form builds bit for bit, eagerly and from three graph
    # replays, for FP64, int8 and int8 with its contraction and density screened at 2^-50 and
    # the contraction's kept blocks packed into tasks (every molecule's 1/2 tr(P F) within
    # 1e-6 Eh of the FP64 P-form build)
    import napmo_amd
    from napmo_amd import workloads
    from napmo_amd.device import DeviceArray
    from napmo_amd.system import cext
    mols, grids, bases, Ps = batch64
    for prec in ('fp64', 'int8', 'int8s'):
        plans = [napmo_amd.GridFock(g, b, m['lmax'], prec[:4] if prec == 'int8s' else prec)
                 for g, b, m in zip(grids, bases, mols)]
        if prec == 'int8s':
            for f in plans:
                f.set_screen(50)
                f.set_density_screen(50)
                f.set_int8_pack(True)   # packed tasks, kept across builds and graph replays
        rng = np.random.default_rng(5)
        Cs = [workloads.occupied_orbitals(rng, f.overlap(), m['nocc']) for f, m in zip(plans, mols)]
        assert all(np.allclose(2.0 * C @ C.T, P, atol=1e-12) for C, P in zip(Cs, Ps))
        single = [f.build_orbitals(C) for f, C in zip(plans, Cs)]
        dEt = np.array([abs(x[1] - r[1]) for x, r in zip(single, fp64_builds)])
        print('%s orbital form: max |dE_trace| vs fp64 %.3e Eh' % (prec, dEt.max()))
        assert dEt.max() <= GATE_EH
        dC = [DeviceArray.from_host(C) for C in Cs]
        dF = [DeviceArray(P.shape) for P in Ps]
        dE = [DeviceArray((2,)) for _ in Ps]
        nocc = [C.shape[1] for C in Cs]
        batch = napmo_amd.BatchFock(plans, nlanes=4)
        batch.build_orbitals(_ptr(dC), nocc, 2.0, _ptr(dF), _ptr(dE))
        batch.synchronize()
        for k, (F, E, Eg) in enumerate(single):
            Eb = dE[k].to_host()
            assert np.array_equal(dF[k].to_host(), F) and Eb[0] == E and Eb[1] == Eg, (prec, k)
        batch.capture_orbitals(_ptr(dC), nocc, 2.0, _ptr(dF), _ptr(dE))
        for it in range(3):
            for a, P in zip(dF, Ps):  # outputs cleared, so every replay rewrites them
                Z = np.zeros_like(P)
                cext.check(cext.napmo_memcpy(a.ptr, Z.ctypes.data, Z.nbytes, 1))
            batch.replay()
            batch.synchronize()
            for k, (F, E, Eg) in enumerate(single):
                Eb = dE[k].to_host()
                assert np.array_equal(dF[k].to_host(), F) and Eb[0] == E and Eb[1] == Eg, (prec, it, k)
        print('%s orbital form: 64 single builds = eager batch = 3 graph replays' % prec)
        with pytest.raises(cext.NapmoError):
            napmo_amd.BatchFock(plans, nlanes=0).build_orbitals(_ptr(dC), nocc, 2.0, _ptr(dF), _ptr(dE))
        del batch, plans
