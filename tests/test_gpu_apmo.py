"""GPU parity of the multi-species (APMO) Coulomb build: electrons and quantum protons,
each species on its own Becke grid, coupled through their Poisson potentials."""
import numpy as np
import pytest

from oracle import cext as ocx
from oracle import core
from tests import systems

pytestmark = pytest.mark.gpu


def _evaluate_c(grid, dens, lmax, points):
    """Potential of dens on grid at points: oracle Poisson, C-oracle synthesis."""
    U = core.poisson_solver(grid, dens, lmax)
    out = np.zeros(len(points))
    for k in range(grid.ncenter):
        ocx.eval_decomposition(U[k], grid.origin[k], points, out)
    return out


def _small_system(seed):
    rng = np.random.default_rng(seed)
    atoms = systems.random_molecule(rng, 3, 4.0, elements=('C', 'N', 'O'), min_sep=1.4)
    hs = [('H', tuple(float(c) for c in rng.uniform(0, 4.0, 3))) for _ in range(2)]
    e_atoms = atoms + hs
    e_shells = systems.random_shells(rng, len(e_atoms), 3)
    from napmo_amd.workloads import even_tempered_shells
    n_shells = [even_tempered_shells(nexp=3) for _ in hs]
    return e_atoms, e_shells, hs, n_shells


def _species_pair(e_atoms, e_shells, n_atoms, n_shells, e_grid, n_grid, shared):
    """(gpu species list, oracle species dicts) with P from occupied orbitals."""
    import napmo_amd
    ge, be = systems.gpu_system(e_atoms, e_shells, *e_grid)
    oe, obe = systems.oracle_system(e_atoms, e_shells, *e_grid)
    if shared:
        gn, on = ge, oe
        bn = napmo_amd.make_basis(n_atoms, shells_per_atom=n_shells)
        obn = core.Basis()
        for (s, o), sh in zip(n_atoms, n_shells):
            for shell in sh:
                obn.add_shell(o, shell['l'], shell['prim'], shell['cont'])
    else:
        gn, bn = systems.gpu_system(n_atoms, n_shells, *n_grid)
        on, obn = systems.oracle_system(n_atoms, n_shells, *n_grid)
    phi_e = ocx.basis_compute(obe, oe.points)
    phi_n = ocx.basis_compute(obn, on.points)
    rng = np.random.default_rng(7)
    Se = core.coupling_matrix(phi_e, np.ones(oe.size), oe.becke_weights, oe.weights)
    Sn = core.coupling_matrix(phi_n, np.ones(on.size), on.becke_weights, on.weights)
    Ce = systems.occupied_orbitals(rng, Se, 3)
    Cn = systems.occupied_orbitals(rng, Sn, len(n_atoms))
    Pe, Pn = 2.0 * Ce @ Ce.T, Cn @ Cn.T
    lm_e, lm_n = 8, 6
    gpu = [(ge, be, -1.0, lm_e), (gn, bn, 1.0, lm_n)]
    ora = [dict(grid=oe, phi=phi_e, P=Pe, charge=-1.0, lmax=lm_e),
           dict(grid=on, phi=phi_n, P=Pn, charge=1.0, lmax=lm_n)]
    return gpu, ora, [Pe, Pn]


@pytest.mark.parametrize('shared', [False, True])
def test_apmo_matches_oracle(shared):
    # electrons on 5 atoms (30 x 110), protons on 2 centres (20 x 50) or on the electronic
    # grid (the reference's arrangement); FP64 throughout
    import napmo_amd
    e_atoms, e_shells, hs, n_shells = _small_system(41)
    gpu, ora, Ps = _species_pair(e_atoms, e_shells, hs, n_shells, (30, 110), (20, 50), shared)
    plan = napmo_amd.APMOFock(gpu, 'fp64')
    Fs, E = plan.build(Ps)
    ref = core.apmo_coulomb_build(ora, evaluate=_evaluate_c)
    for s, (F, (Fr, Etr, Egr, Esr)) in enumerate(zip(Fs, ref)):
        dF = np.max(np.abs(F - Fr))
        print('apmo shared=%s species %d n=%d max|dF| %.3e E %s ref %.12f %.12f %.12f' %
              (shared, s, F.shape[0], dF, E[s], Etr, Egr, Esr))
        assert dF < 1e-10 * np.abs(Fr).max()
        assert np.array_equal(F, F.T)
        assert abs(E[s, 0] - Etr) < 1e-10 * abs(Etr)
        assert abs(E[s, 1] - Egr) < 1e-10 * abs(Egr)
        assert abs(E[s, 2] - Esr) < 1e-10 * abs(Esr)
        assert abs(E[s, 3] - (Egr - Esr)) < 1e-10 * abs(Egr)
    # attraction: the two coupling energies are negative and equal in the continuum
    # (1/2 q_e q_n int rho_e V[rho_n] = 1/2 q_n q_e int rho_n V[rho_e]); on these coarse
    # grids they differ by the discretisation (oracle: 3.7e-3 relative on separate grids,
    # 8e-4 on the shared one)
    assert E[0, 3] < 0 and E[1, 3] < 0
    assert abs(E[0, 3] - E[1, 3]) < 1e-2 * abs(E[0, 3])
    # repeated builds are deterministic
    Fs2, E2 = plan.build(Ps)
    assert all(np.array_equal(a, b) for a, b in zip(Fs, Fs2)) and np.array_equal(E, E2)


def test_apmo_full_size_properties():
    # 4 heavy atoms + 4 H: electrons on 8 x 100 x 302 points (241,600), 120 functions;
    # 4 quantum protons on 4 x 60 x 194 (46,560), 320 even-tempered functions
    import napmo_amd
    from napmo_amd import workloads
    sysd = workloads.apmo_system()
    e, n = sysd['electron'], sysd['proton']
    ge, be = systems.gpu_system(e['centres'], e['shells'], e['nrad'], e['nang'])
    gn, bn = systems.gpu_system(n['centres'], n['shells'], n['nrad'], n['nang'])
    print('apmo N_e', ge.size, 'n_e', be['nbasis'], 'N_n', gn.size, 'n_n', bn['nbasis'])
    assert ge.size == 241600 and gn.size == 46560
    assert be['nbasis'] == 120 and bn['nbasis'] == 320
    fe = napmo_amd.GridFock(ge, be, e['lmax'], 'fp64')
    fn = napmo_amd.GridFock(gn, bn, n['lmax'], 'fp64')
    rng = np.random.default_rng(3)
    Ce = workloads.occupied_orbitals(rng, fe.overlap(), e['nocc'])
    Cn = workloads.occupied_orbitals(rng, fn.overlap(), n['nocc'])
    Pe, Pn = e['eta'] * Ce @ Ce.T, n['eta'] * Cn @ Cn.T
    del fe, fn
    species = [(ge, be, e['charge'], e['lmax']), (gn, bn, n['charge'], n['lmax'])]
    plan = napmo_amd.APMOFock(species, 'fp64')
    Fs, E = plan.build([Pe, Pn])
    print('apmo E_e', E[0], 'E_n', E[1])
    for s in range(2):
        assert np.array_equal(Fs[s], Fs[s].T)
        assert abs(E[s, 0] - E[s, 1]) < 1e-9 * abs(E[s, 1])
    assert E[0, 3] < 0 and E[1, 3] < 0
    # reciprocity of the e-n coupling across the two grids (discretisation level)
    rel = abs(E[0, 3] - E[1, 3]) / abs(E[0, 3])
    print('apmo coupling e %.10f n %.10f rel %.2e' % (E[0, 3], E[1, 3], rel))
    assert rel < 1e-2
    # mixed modes against FP64, per species, on the energy of the F each mode returns
    # (E[s, 0] = 1/2 tr(P F)); E[s, 1] comes from the FP64 rho and v in every mode
    F8 = None
    for prec in ('int8', 'int8s', 'bf16x3'):
        mix = napmo_amd.APMOFock(species, prec[:4] if prec == 'int8s' else prec)
        if prec == 'int8s':
            mix.set_screen(50)    # blocks bounded by 2^-50 per element and chunk skipped
        Fm, Em = mix.build([Pe, Pn])
        if prec == 'int8':
            F8 = Fm
        for s in range(2):
            dF = np.max(np.abs(Fm[s] - Fs[s]))
            dEt = Em[s, 0] - E[s, 0]
            print('apmo %s species %d max|dF| %.3e dE_grid %.3e dE_trace %.3e' %
                  (prec, s, dF, Em[s, 1] - E[s, 1], dEt))
            assert abs(Em[s, 1] - E[s, 1]) < 1e-9 * abs(E[s, 1])
            if prec == 'int8s':
                # within the stated bound (chunks x 2^-50 per element) of the unscreened
                # contraction, up to the rounding of the reordered operands
                assert np.max(np.abs(Fm[s] - F8[s])) <= 1e-9 * np.abs(F8[s]).max()
            if prec.startswith('int8'):
                # exact int32 accumulation of 28-bit digit slices: within 1e-6 Eh
                assert abs(dEt) <= 1e-6
                assert dF < 1e-7 * np.abs(Fs[s]).max()
            else:
                # BF16 x 3 with FP32 accumulation: about 4e-5 Eh here, outside 1e-6 Eh
                assert abs(dEt) < 1e-3
                assert dF < 2e-5 * np.abs(Fs[s]).max()
        del mix


@pytest.mark.parametrize('shared', [False, True])
def test_apmo_one_launch_matches_per_species_build(shared):
    # every phase for both species as one launch (own and cross-species pairs in one pair
    # list per lmax, grouped contractions): F and all four energies per species equal the
    # per-species build bit for bit (the same operations in the same order), on separate
    # and on shared grids
    import napmo_amd
    e_atoms, e_shells, hs, n_shells = _small_system(43)
    gpu, ora, Ps = _species_pair(e_atoms, e_shells, hs, n_shells, (30, 110), (20, 50), shared)
    ref = napmo_amd.APMOFock(gpu, 'fp64')
    Fr, Er = ref.build(Ps)
    one = napmo_amd.APMOFock(gpu, 'fp64', one_launch=True)
    Fo, Eo = one.build(Ps)
    for s in range(2):
        dF = np.abs(Fo[s] - Fr[s]).max() / np.abs(Fr[s]).max()
        dE = np.abs(Eo[s] - Er[s]).max() / np.abs(Er[s]).max()
        print('shared %s species %d: rel max|dF| %.2e, rel max|dE| %.2e, E %s' % (shared, s, dF, dE, Eo[s]))
        assert np.array_equal(Fo[s], Fr[s]) and np.array_equal(Eo[s], Er[s])
    Fo2, Eo2 = one.build(Ps)
    assert all(np.array_equal(a, b) for a, b in zip(Fo, Fo2)) and np.array_equal(Eo, Eo2)
    # and back to the per-species path
    one.set_one_launch(False)
    Fb, Eb = one.build(Ps)
    assert all(np.array_equal(a, b) for a, b in zip(Fb, Fr)) and np.array_equal(Eb, Er)
    with pytest.raises(Exception, match='FP64 plans'):
        napmo_amd.APMOFock(gpu, 'int8', one_launch=True)


def test_apmo_one_launch_graph_replays_match_eager():
    # the one-launch build captured once into a HIP graph; replays on changing densities,
    # with eager builds and host synchronisation between them, equal eager builds bit for bit
    import napmo_amd
    from napmo_amd.device import DeviceArray
    from napmo_amd.system import cext
    e_atoms, e_shells, hs, n_shells = _small_system(44)
    gpu, ora, Ps = _species_pair(e_atoms, e_shells, hs, n_shells, (30, 110), (20, 50), False)
    plan = napmo_amd.APMOFock(gpu, 'fp64', one_launch=True)
    dP = [DeviceArray.from_host(P) for P in Ps]
    dF = [DeviceArray(P.shape) for P in Ps]
    dE = DeviceArray((8,))
    plan.capture([a.ptr for a in dP], [a.ptr for a in dF], dE.ptr)
    for it in range(3):
        scale = 1.0 + 0.3 * it
        for a, P in zip(dP, Ps):
            Pi = np.ascontiguousarray(P * scale)
            cext.check(cext.napmo_memcpy(a.ptr, Pi.ctypes.data, Pi.nbytes, 1))
        plan.replay()
        plan.synchronize()
        got = [a.to_host() for a in dF], dE.to_host()
        Fe, Ee = plan.build([P * scale for P in Ps])
        assert all(np.array_equal(a, b) for a, b in zip(got[0], Fe)), it
        assert np.array_equal(got[1], Ee.reshape(-1)), it
        print('replay', it, 'E', Ee[:, 1])
    with pytest.raises(Exception, match='one-launch'):
        napmo_amd.APMOFock(gpu, 'fp64').replay()


def test_apmo_one_launch_three_species():
    # electrons and two proton species on their own grids: every species couples to the two
    # others; the one-launch build equals the per-species build bit for bit
    import napmo_amd
    from napmo_amd.workloads import even_tempered_shells
    e_atoms, e_shells, hs, n_shells = _small_system(45)
    ge, be = systems.gpu_system(e_atoms, e_shells, 30, 110)
    species, Ps = [(ge, be, -1.0, 8)], []
    for h in hs:
        gn, bn = systems.gpu_system([h], [even_tempered_shells(nexp=3)], 20, 50)
        species.append((gn, bn, 1.0, 6))
    rng = np.random.default_rng(9)
    for g, b, q, lm in species:
        S = napmo_amd.GridFock(g, b, lm, 'fp64').overlap()
        C = systems.occupied_orbitals(rng, S, 2 if q < 0 else 1)
        Ps.append((2.0 if q < 0 else 1.0) * C @ C.T)
    Fr, Er = napmo_amd.APMOFock(species, 'fp64').build(Ps)
    Fo, Eo = napmo_amd.APMOFock(species, 'fp64', one_launch=True).build(Ps)
    print('three species: E per species', Eo[:, 1], 'coupling', Eo[:, 3])
    for s in range(3):
        assert np.array_equal(Fo[s], Fr[s]) and np.array_equal(Eo[s], Er[s]), s
