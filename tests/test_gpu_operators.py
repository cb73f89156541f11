"""GPU grid one-electron operators: the kinetic operator (projection, radial finite
differences, spline, synthesis) and the point-charge potential, against the oracle."""
import numpy as np
import pytest

from oracle import core
from tests import systems

pytestmark = pytest.mark.gpu


def test_kinetic_matches_oracle_and_closed_form():
    import napmo_amd
    # three centres, 40 x 110 points each, lmax 6; psi a sum of Gaussians on the centres
    atoms = [('O', (0.0, 0.0, 0.0)), ('H', (1.1, 0.9, 0.0)), ('H', (-1.2, 0.7, 0.2))]
    g, _ = systems.gpu_system(atoms, [[]] * 3, 40, 110)
    o, _ = systems.oracle_system(atoms, [[]] * 3, 40, 110)
    psi = sum(np.exp(-0.9 * ((o.points - np.array(c)) ** 2).sum(1)) for _, c in atoms)
    T = napmo_amd.compute_kinetic(g, psi, 6)
    ref = core.compute_kinetic(o, psi, 6)
    print('kinetic max|dT| %.3e max|T| %.3e' % (np.abs(T - ref).max(), np.abs(ref).max()))
    assert np.max(np.abs(T - ref)) < 1e-11 * np.abs(ref).max()
    # one centre, fine radial grid: -1/2 laplacian of a Gaussian in closed form
    g1, _ = systems.gpu_system([('H', (0.0, 0.0, 0.0))], [[]], 120, 50)
    r = np.sqrt((g1.points ** 2).sum(1))
    psi1 = np.exp(-0.8 * r * r)
    T1 = napmo_amd.compute_kinetic(g1, psi1, 2)
    ex = -0.5 * (4 * 0.64 * r * r - 4.8) * psi1
    assert np.max(np.abs(T1 - ex)) < 2e-4


def test_nuclear_matches_oracle():
    import napmo_amd
    atoms = systems.random_molecule(np.random.default_rng(5), 5, 5.0)
    g, _ = systems.gpu_system(atoms, [[]] * 5, 30, 50)
    charges = [{'atomic_number': {'H': 1, 'C': 6, 'N': 7, 'O': 8}[s], 'origin': np.array(o)} for s, o in atoms]
    V = napmo_amd.compute_nuclear(g, charges)
    ref = core.compute_nuclear(g.points, [(c['atomic_number'], c['origin']) for c in charges])
    assert np.max(np.abs(V - ref) / np.abs(ref)) < 1e-14
