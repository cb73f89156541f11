"""Screened orbital-form density on the GPU (NAPMO_TUNE_DENSITY_SCREEN, the streamed kernel):

* rho against the oracle's restatement with the same threshold (oracle.core.
  density_orbitals_screened: same function order, block exponents and left-out blocks) to
  FP64 rounding, and against the unscreened density within the stated bound, on a five-atom
  system where part of the (tile, group) products is left out;
* the 16-atom system with both screenings (density and int8 contraction at 2^-50): the grid
  and trace energies within 1e-6 Eh of the FP64 build."""
import numpy as np
import pytest

from oracle import core
from tests import systems

pytestmark = pytest.mark.gpu


def _bound(X, G, S, eta):
    return eta * (2 * np.abs(X).sum(1) * G * 2.0 ** -S + X.shape[1] * (G * 2.0 ** -S) ** 2)


def test_screened_density_matches_the_oracle_restatement():
    import napmo_amd
    rng = np.random.default_rng(71)
    atoms = systems.random_molecule(rng, 5, 8.0, min_sep=1.2)
    shells = systems.random_shells(rng, 5, 7)               # 70 functions: 5 groups of 16
    g, b = systems.gpu_system(atoms, shells, 24, 110)
    fock = napmo_amd.GridFock(g, b, 8, 'fp64')
    C = systems.occupied_orbitals(np.random.default_rng(5), fock.overlap(), 12)
    phi = b.compute(g.points)
    X = phi @ C
    full = 2.0 * (X ** 2).sum(1)
    G = -(-b['nbasis'] // 16)
    _, _, _, rho0, _ = fock.build_orbitals(C, 2.0, want_grid=True)
    assert np.abs(rho0 - full).max() <= 1e-12 * full.max()
    for S in (20, 40):
        fock.set_density_screen(S)
        _, _, _, rho, _ = fock.build_orbitals(C, 2.0, want_grid=True)
        ro, order, kept = core.density_orbitals_screened(phi, C, 2.0, S)
        d = np.abs(rho - full)
        print('S %d: %.1f%% of the tile groups kept; max|rho - oracle| %.3e, max|rho - unscreened| %.3e'
              % (S, 100 * kept, np.abs(rho - ro).max(), d.max()))
        assert kept < 1.0
        assert np.abs(rho - ro).max() <= 1e-12 * full.max()
        assert (d <= _bound(X, G, S, 2.0) + 1e-12 * full.max()).all()
    fock.set_density_screen(0)
    _, _, _, rho, _ = fock.build_orbitals(C, 2.0, want_grid=True)
    assert np.array_equal(rho, rho0)


def test_both_screenings_on_the_16_atom_system():
    import napmo_amd
    atoms, shells, nr, na, lmax = systems.chon16_system()
    g, b = systems.gpu_system(atoms, shells, nr, na)
    f64 = napmo_amd.GridFock(g, b, lmax, 'fp64')
    C = systems.occupied_orbitals(np.random.default_rng(2), f64.overlap(), 40)
    F1, E1, Eg1 = f64.build_orbitals(C, 2.0)
    del f64
    i8 = napmo_amd.GridFock(g, b, lmax, 'int8')
    i8.set_screen(50)
    i8.set_density_screen(50)
    F2, E2, Eg2 = i8.build_orbitals(C, 2.0)
    print('chon16 int8, both screenings at 2^-50: dE_trace %.3e dE_grid %.3e max|dF| %.3e' %
          (E2 - E1, Eg2 - Eg1, np.abs(F2 - F1).max()))
    assert abs(E2 - E1) <= 1e-6 and abs(Eg2 - Eg1) <= 1e-6
    assert np.abs(F2 - F1).max() <= 1e-7 * np.abs(F1).max()


def test_screening_edge_cases_small_basis_one_orbital_aggressive_threshold():
    # H2 with a cc-pVDZ-size basis (10 functions: one partial block / group), one orbital and
    # the most aggressive thresholds: both screened paths still equal their oracle
    # restatements, and the contraction's left-out blocks are counted as the oracle counts them
    import napmo_amd
    g, b = systems.gpu_system(systems.H2_ATOMS, systems.named_shells(systems.H2_ATOMS, 'CC-PVDZ'), 40, 110)
    phi = b.compute(g.points)
    wbw = g.weights * g.becke_weights
    C = np.ascontiguousarray(np.linalg.qr(np.random.default_rng(8).normal(size=(b['nbasis'], 1)))[0]) * 0.7
    f = napmo_amd.GridFock(g, b, 6, 'int8')
    for S in (1, 8):
        f.set_density_screen(S)
        _, _, _, rho, _ = f.build_orbitals(C, 2.0, want_grid=True)
        ro, _, kept = core.density_orbitals_screened(phi, C, 2.0, S)
        print('H2 density S %d: kept %.3f, max|rho - oracle| %.3e' % (S, kept, np.abs(rho - ro).max()))
        assert np.abs(rho - ro).max() <= 1e-12 * max(ro.max(), 1e-300)
    f.set_density_screen(0)
    v = 1.0 / (0.5 + (g.points ** 2).sum(1))
    chunk = 256
    from napmo_amd.system import cext
    cext.check(cext.napmo_fock_plan_set_tuning(f._this, 7, chunk))
    order = core.int8_compact_order(phi, wbw, chunk)
    for S in (1, 8, 30):
        f.set_screen(S)
        F = f.contract(v)
        st = {}
        Fo = core.int8_fock(phi, wbw, v, chunk, screen=S, order=order, stats=st)
        print('H2 int8 S %d: %d of %d chunk blocks left out' % (S, st['skipped'], st['kept'] + st['skipped']))
        assert np.array_equal(F, Fo)
        assert f.screen_stats() == (st['kept'], st['kept'] + st['skipped'])
