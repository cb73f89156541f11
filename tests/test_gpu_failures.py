"""Forced failures through the status-returning entry points (napmo_*): each returns a
non-zero status with a message (cext raises NapmoError with it), no kernel is left running on
bad data, and the plan stays usable afterwards. The conditions are set from the host; every
kernel that runs exits normally.

* a non-finite value in P (or in C, or in a potential): reported with its position, before
  anything is uploaded;
* a finite P whose density overflows (P = 1e300 I): the build's energy is not finite and is
  reported instead of returned;
* a singular radial Poisson system: a Chebyshev transform of radius 0 puts every radial
  point at r = 0, where the radial ODE divides by r (reported when the plan is made, and by
  the standalone Coulomb solve);
* a non-finite value in one species' P of an APMO build."""
import numpy as np
import pytest

from tests import systems

pytestmark = pytest.mark.gpu


def _h2_plan(prec='fp64'):
    import napmo_amd
    atoms = systems.H2_ATOMS
    g, b = systems.gpu_system(atoms, systems.named_shells(atoms, 'CC-PVDZ'), 40, 110)
    f = napmo_amd.GridFock(g, b, 6, prec)
    P = systems.density_matrix(np.random.default_rng(7), f.overlap(), 1)
    return g, b, f, P


def test_non_finite_density_matrix_is_reported_and_plan_survives():
    from napmo_amd.system.cext import NapmoError
    for prec in ('fp64', 'int8'):
        g, b, f, P = _h2_plan(prec)
        F0, E0, Eg0 = f.build(P)
        bad = P.copy()
        bad[3, 5] = np.nan
        with pytest.raises(NapmoError, match=r'P holds a non-finite value at row 3, column 5'):
            f.build(bad)
        bad[3, 5] = np.inf
        with pytest.raises(NapmoError, match=r'non-finite'):
            f.build(bad)
        F1, E1, Eg1 = f.build(P)           # the plan is unchanged by the refused calls
        assert np.array_equal(F1, F0) and E1 == E0 and Eg1 == Eg0
        print(prec, 'non-finite P refused; E', E0)


def test_non_finite_potential_is_reported():
    from napmo_amd.system.cext import NapmoError
    g, b, f, P = _h2_plan()
    v = np.ones(g.size)
    v[11] = np.nan
    with pytest.raises(NapmoError, match=r'v holds a non-finite value at row 11'):
        f.contract(v)


def test_overflowing_density_gives_an_error_not_an_energy():
    from napmo_amd.system.cext import NapmoError
    g, b, f, P = _h2_plan()
    big = 1e300 * np.eye(P.shape[0])
    with pytest.raises(NapmoError, match=r'non-finite energy'):
        f.build(big)
    F1, E1, Eg1 = f.build(P)
    assert np.isfinite(E1) and np.isfinite(F1).all()


def test_singular_radial_system_is_reported():
    import napmo_amd
    from napmo_amd.grids.radial_transform import ChebyshevRadialTransform
    from napmo_amd.system.cext import NapmoError
    atoms = systems.H2_ATOMS
    grid = napmo_amd.BeckeGrid(napmo_amd.make_species(atoms), 20, 110,
                               rtransform=ChebyshevRadialTransform(0.0, 20))
    basis = napmo_amd.make_basis(atoms, shells_per_atom=systems.named_shells(atoms, 'CC-PVDZ'))
    with pytest.raises(NapmoError, match=r'singular radial Poisson system: radial point 0 at r = 0'):
        napmo_amd.GridFock(grid, basis, 4, 'fp64')
    with pytest.raises(NapmoError, match=r'singular radial Poisson system'):
        napmo_amd.compute_coulomb(grid, np.ones(grid.size), 4)


def test_apmo_non_finite_species_density_is_reported():
    import napmo_amd
    from napmo_amd.system.cext import NapmoError
    g, b, f, P = _h2_plan()
    species = [(g, b, -1.0, 6), (g, b, -1.0, 6)]
    plan = napmo_amd.APMOFock(species, 'fp64')
    Fs, E = plan.build([P, P])
    bad = P.copy()
    bad[0, 0] = np.nan
    with pytest.raises(NapmoError, match=r'APMO build: P of species 1 holds a non-finite value at row 0, column 0'):
        plan.build([P, bad])
    Fs2, E2 = plan.build([P, P])
    assert np.array_equal(E2, E)
