"""int8 digit-sliced grid Fock contraction (NAPMO_PREC_INT8) on the GPU:
  * bit for bit against the oracle's exact restatement of its integer arithmetic
    (oracle.core.int8_fock) on the same Phi, weights and potential, over several chunks,
    stages, row groups and a grid with negative Lebedev weights;
  * its accuracy against the FP64 build: |dE_trace| = |1/2 tr(P (F_int8 - F_fp64))| <= 1e-6 Eh
    on the 16-atom 472,000-point system (the stated mixed-precision tolerance), and far below
    it on the small systems;
  * screening (NAPMO_TUNE_INT8_SCREEN): bit for bit against the oracle's restatement with the
    same threshold and function order, and on the 16-atom system within the stated element
    bound of the unscreened contraction and the 1e-6 Eh tolerance of the FP64 build."""
import numpy as np
import pytest

from oracle import core
from tests import systems

pytestmark = pytest.mark.gpu

TUNE_INT8_CHUNK = 7
TUNE_INT8_SCREEN = 11


def _set_screen(fock, S):
    from napmo_amd.system import cext
    cext.check(cext.napmo_fock_plan_set_tuning(fock._this, TUNE_INT8_SCREEN, S))


def _plan_chunk(fock, chunk):
    import ctypes
    from napmo_amd.system import cext
    cext.check(cext.napmo_fock_plan_set_tuning(fock._this, TUNE_INT8_CHUNK, chunk))
    v = ctypes.c_long()
    cext.check(cext.napmo_fock_plan_get_tuning(fock._this, TUNE_INT8_CHUNK, ctypes.byref(v)))
    return v.value


@pytest.mark.parametrize('case', ['five_atoms', 'negative_weights'])
def test_int8_contraction_is_bit_exact_with_the_oracle_restatement(case):
    import napmo_amd
    rng = np.random.default_rng(71)
    if case == 'five_atoms':
        atoms = systems.random_molecule(rng, 5, 5.0, min_sep=1.2)
        shells = systems.random_shells(rng, 5, 7)            # 70 functions: 5 blocks, two groups
        g, b = systems.gpu_system(atoms, shells, 24, 110)
    else:
        atoms = systems.random_molecule(rng, 3, 4.0, min_sep=1.2)
        shells = systems.random_shells(rng, 3, 4)
        g, b = systems.gpu_system(atoms, shells, None, None, sizes=[(20, 230), (18, 266), (16, 74)])
        assert (g.weights < 0).any()
    fock = napmo_amd.GridFock(g, b, 8, 'int8')
    chunk = _plan_chunk(fock, 320)  # 5 stages per chunk, many chunks
    assert chunk == 320
    phi = b.compute(g.points)
    wbw = g.weights * g.becke_weights
    v = 1.0 / (0.5 + ((g.points - g.origin[0]) ** 2).sum(1)) - 0.3 * np.exp(-((g.points - g.origin[1]) ** 2).sum(1))
    F = fock.contract(v)
    Fo = core.int8_fock(phi, wbw, v, chunk)
    Fr = core.coupling_matrix(phi, v, g.becke_weights, g.weights)
    print('%s: n %d N %d max|F - oracle| %.3e max|F - fp64| %.3e (max|F| %.3e)' % (
        case, b['nbasis'], g.size, np.abs(F - Fo).max(), np.abs(F - Fr).max(), np.abs(Fr).max()))
    assert np.array_equal(F, Fo)
    # 28-bit operands: ~1e-8 of max|F| here; the sign-changing potential on the grid with
    # negative Lebedev weights cancels more (measured 2.8e-8)
    tol = 1e-8 if case == 'five_atoms' else 1e-7
    assert np.abs(F - Fr).max() <= tol * np.abs(Fr).max()
    # the default chunking gives the same accuracy
    _plan_chunk(fock, 0)
    assert np.abs(fock.contract(v) - Fr).max() <= tol * np.abs(Fr).max()


def test_int8_build_matches_fp64_build_small():
    import napmo_amd
    for atoms, shells, nr, na, lmax in (
            (systems.H2_ATOMS, systems.named_shells(systems.H2_ATOMS, 'CC-PVDZ'), 100, 110, 10),
            (systems.random_molecule(np.random.default_rng(31), 5, 5.0, min_sep=1.2),
             systems.random_shells(np.random.default_rng(31), 5, 6), 30, 110, 8)):
        g, b = systems.gpu_system(atoms, shells, nr, na)
        f64 = napmo_amd.GridFock(g, b, lmax, 'fp64')
        i8 = napmo_amd.GridFock(g, b, lmax, 'int8')
        P = systems.density_matrix(np.random.default_rng(3), f64.overlap(), 2)
        F1, E1, Eg1 = f64.build(P)
        F2, E2, Eg2 = i8.build(P)
        print('int8 n=%d dE_trace %.3e max|dF| %.3e (max|F| %.3e)' % (
            b['nbasis'], E2 - E1, np.abs(F2 - F1).max(), np.abs(F1).max()))
        assert abs(E2 - E1) <= 1e-7    # the stated mixed-mode tolerance is 1e-6 Eh
        assert abs(Eg2 - Eg1) <= 1e-12 * abs(Eg1)
        assert np.abs(F2 - F1).max() <= 1e-8 * np.abs(F1).max()
        assert np.array_equal(F2, F2.T)


def test_int8_build_meets_the_stated_tolerance_on_the_16_atom_system():
    # 16 atoms, 472,000 points, 300 functions, lmax 12: |1/2 tr(P F_int8) - 1/2 tr(P F_fp64)|
    # <= 1e-6 Eh, with the orbital-form density (the benchmarked build) and the P form
    import napmo_amd
    atoms, shells, nr, na, lmax = systems.chon16_system()
    g, b = systems.gpu_system(atoms, shells, nr, na)
    f64 = napmo_amd.GridFock(g, b, lmax, 'fp64')
    C = systems.occupied_orbitals(np.random.default_rng(2), f64.overlap(), 40)
    F1, E1, Eg1 = f64.build_orbitals(C, 2.0)
    del f64
    i8 = napmo_amd.GridFock(g, b, lmax, 'int8')
    F2, E2, Eg2 = i8.build_orbitals(C, 2.0)
    F3, E3, Eg3 = i8.build(2.0 * C @ C.T)
    print('chon16 int8: E %.10f dE_trace %.3e (P form %.3e) max|dF| %.3e max|F| %.3e' % (
        E1, E2 - E1, E3 - E1, np.abs(F2 - F1).max(), np.abs(F1).max()))
    assert abs(E2 - E1) <= 1e-6
    assert abs(E3 - E1) <= 1e-6
    assert np.abs(F2 - F1).max() <= 1e-7 * np.abs(F1).max()
    assert np.array_equal(F2, F2.T)


def test_int8_screening_is_bit_exact_with_the_oracle_restatement():
    # five atoms (70 functions, 5 blocks, 42 chunks of 320 points): at 2^-30 and 2^-50 the
    # oracle leaves out 147 and 120 of the 630 chunk blocks; 2^-1000 leaves out none but the
    # function order changes
    import napmo_amd
    rng = np.random.default_rng(71)
    atoms = systems.random_molecule(rng, 5, 5.0, min_sep=1.2)
    shells = systems.random_shells(rng, 5, 7)
    g, b = systems.gpu_system(atoms, shells, 24, 110)
    fock = napmo_amd.GridFock(g, b, 8, 'int8')
    chunk = _plan_chunk(fock, 320)
    phi = b.compute(g.points)
    wbw = g.weights * g.becke_weights
    v = 1.0 / (0.5 + ((g.points - g.origin[0]) ** 2).sum(1)) - 0.3 * np.exp(-((g.points - g.origin[1]) ** 2).sum(1))
    F_plain = fock.contract(v)
    order = core.int8_compact_order(phi, wbw, chunk)
    nchunk = -(-g.size // chunk)
    for S in (1000, 50, 30):
        _set_screen(fock, S)
        F = fock.contract(v)
        st = {}
        Fo = core.int8_fock(phi, wbw, v, chunk, screen=S, order=order, stats=st)
        d = np.abs(F - F_plain).max()
        print('screen 2^-%d: %d of %d chunk blocks left out; max|F - oracle| %.3e, max|F - unscreened| %.3e' % (
            S, st['skipped'], st['kept'] + st['skipped'], np.abs(F - Fo).max(), d))
        assert np.array_equal(F, Fo)
        assert fock.screen_stats() == (st['kept'], st['kept'] + st['skipped'])
        assert d <= nchunk * 2.0 ** -min(S, 1000) + 1e-9 * np.abs(F_plain).max()
    _set_screen(fock, 0)
    assert np.array_equal(fock.contract(v), F_plain)
    kept, total = fock.screen_stats()
    assert kept == total
    fock.set_int8_pack(True)    # every block packed per chunk: the unscreened contraction
    assert np.array_equal(fock.contract(v), F_plain)
    fock.set_int8_pack(False)


def test_int8_screening_on_the_16_atom_system():
    # 16 atoms, 472,000 points, 300 functions: the screened build (2^-50 per element and chunk)
    # within the element bound of the unscreened one and within 1e-6 Eh of the FP64 build
    import napmo_amd
    atoms, shells, nr, na, lmax = systems.chon16_system()
    g, b = systems.gpu_system(atoms, shells, nr, na)
    f64 = napmo_amd.GridFock(g, b, lmax, 'fp64')
    C = systems.occupied_orbitals(np.random.default_rng(2), f64.overlap(), 40)
    F1, E1, Eg1 = f64.build_orbitals(C, 2.0)
    del f64
    i8 = napmo_amd.GridFock(g, b, lmax, 'int8')
    F2, E2, Eg2 = i8.build_orbitals(C, 2.0)
    _set_screen(i8, 50)
    F3, E3, Eg3 = i8.build_orbitals(C, 2.0)
    i8.set_int8_pack(True)     # packed tasks: the same F bit for bit
    F4, E4, Eg4 = i8.build_orbitals(C, 2.0)
    assert np.array_equal(F4, F3) and E4 == E3
    i8.set_int8_pack(False)
    import ctypes
    from napmo_amd.system import cext
    chunk = ctypes.c_long()
    cext.check(cext.napmo_fock_plan_get_tuning(i8._this, TUNE_INT8_CHUNK, ctypes.byref(chunk)))
    nchunk = -(-g.size // chunk.value)
    print('chon16 int8 screened 2^-50: dE_trace %.3e (unscreened %.3e), max|F - unscreened| %.3e, '
          'max|F - fp64| %.3e' % (E3 - E1, E2 - E1, np.abs(F3 - F2).max(), np.abs(F3 - F1).max()))
    assert Eg3 == Eg2
    assert abs(E3 - E1) <= 1e-6
    assert np.abs(F3 - F2).max() <= nchunk * 2.0 ** -50 + 1e-9 * np.abs(F2).max()
    assert np.abs(F3 - F1).max() <= 1e-7 * np.abs(F1).max()
    assert np.array_equal(F3, F3.T)


@pytest.mark.parametrize('seed,natom,ncycle', [(11, 2, 4), (12, 4, 8), (13, 6, 12)])
def test_int8_screening_bit_exact_on_random_molecules(seed, natom, ncycle):
    # random molecules of 20-120 functions in a 9 bohr box at thresholds 2^-20 and 2^-50:
    # the screened contraction equals the oracle's restatement bit for bit, and its count of
    # computed blocks equals the oracle's
    import napmo_amd
    rng = np.random.default_rng(seed)
    atoms = systems.random_molecule(rng, natom, 9.0, min_sep=1.4)
    shells = systems.random_shells(rng, natom, ncycle)
    g, b = systems.gpu_system(atoms, shells, 20, 110)
    fock = napmo_amd.GridFock(g, b, 6, 'int8')
    chunk = _plan_chunk(fock, 384)
    phi = b.compute(g.points)
    wbw = g.weights * g.becke_weights
    v = 2.0 / (0.3 + ((g.points - g.origin[0]) ** 2).sum(1)) + 0.1
    order = core.int8_compact_order(phi, wbw, chunk)
    for S in (20, 50):
        fock.set_screen(S)
        F = fock.contract(v)
        st = {}
        Fo = core.int8_fock(phi, wbw, v, chunk, screen=S, order=order, stats=st)
        print('seed %d n %d S %d: %d of %d chunk blocks left out' % (
            seed, b['nbasis'], S, st['skipped'], st['kept'] + st['skipped']))
        assert np.array_equal(F, Fo)
        assert fock.screen_stats() == (st['kept'], st['kept'] + st['skipped'])
        # the kept blocks packed per chunk into tasks: the same contraction bit for bit, on
        # the first build, on a repeat (tasks reused), after the potential grows (tasks formed
        # again) and after it shrinks back (the grown tasks reused, the extra blocks skipped)
        F_big = fock.contract(8.0 * v)
        fock.set_int8_pack(True)
        assert np.array_equal(fock.contract(v), Fo)
        assert np.array_equal(fock.contract(v), Fo)
        assert np.array_equal(fock.contract(8.0 * v), F_big)
        assert np.array_equal(fock.contract(v), Fo)
        fock.set_int8_pack(False)
