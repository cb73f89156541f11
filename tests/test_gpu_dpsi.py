"""GPU Becke-Dickson orbital correction (napmo/wavefunction/ndpsi.py) and the per-orbital
steps of PSIO_BECKE (napmo/wavefunction/psi_optimization.py:175-203) against the oracle,
and on one centre against the exact solution of the radial equations."""
import numpy as np
import pytest

from oracle import cext as ocx
from oracle import core
from tests import systems
from tests.test_oracle import _dpsi_closed_form_inputs

pytestmark = pytest.mark.gpu

ATOMS = [('O', (0.0, 0.0, 0.0)), ('H', (1.1, 0.9, 0.0)), ('N', (-1.2, 0.7, 0.6))]


def _evaluate_c(splines, centre, points):
    out = np.zeros(len(points))
    ocx.eval_decomposition(splines, centre, points, out)
    return out


def _orbital_inputs(o, norb, seed):
    """Smooth orbitals, residuals and applied potentials on the oracle grid's points."""
    rng = np.random.default_rng(seed)
    p = o.points
    d = [np.sqrt(((p - np.array(c)) ** 2).sum(1)) for _, c in ATOMS]
    phi, ri, V = [], [], []
    for k in range(norb):
        cf = rng.uniform(0.3, 1.0, len(ATOMS))
        ph = sum(cf[a] * np.exp(-(0.6 + 0.2 * k) * d[a] ** 2) for a in range(len(ATOMS)))
        ph = ph * (1.0 + 0.3 * p[:, 2] * (k % 2))
        phi.append(ph)
        ri.append(0.05 * ph * np.cos(p[:, 0]) + 0.02 * np.exp(-d[1] ** 2))
        V.append(-sum(z / np.maximum(r, 1e-3) for z, r in zip((8.0, 1.0, 7.0), d)) * ph * 0.1)
    return np.array(phi), np.array(ri), np.array(V)


def test_dpsi_matches_oracle():
    # three centres, 30 x 50 points, lmax 4, two orbitals (electron) and one with m_inv 1/1836
    import napmo_amd
    g, _ = systems.gpu_system(ATOMS, [[]] * 3, 30, 50)
    o, _ = systems.oracle_system(ATOMS, [[]] * 3, 30, 50)
    phi, ri, V = _orbital_inputs(o, 2, 5)
    doi, oi = np.array([0.03, -0.02]), np.array([-0.7, -0.4])
    for mass_inv in (1.0, 1.0 / 1836.15):
        d = napmo_amd.compute_dpsi(g, 4, phi, doi, oi, ri, V, mass_inv)
        for k in range(2):
            ref = core.compute_dpsi(o, 4, phi[k], doi[k], oi[k], ri[k], V[k], mass_inv,
                                    evaluate=_evaluate_c)
            scale = np.abs(ref).max()
            err = np.abs(d[k] - ref).max()
            print('dpsi m_inv %.3e orbital %d max|d| %.3e max|err| %.3e' % (mass_inv, k, scale, err))
            assert err < 1e-10 * scale
        # the reference's one-orbital signature gives the same row
        d1 = napmo_amd.compute_dpsi(g, 4, phi[1], doi[1], oi[1], ri[1], V[1], mass_inv)
        assert np.array_equal(d1, d[1])


def test_dpsi_closed_form_on_gpu():
    import napmo_amd
    g, _ = systems.gpu_system([('O', (0.0, 0.0, 0.0))], [[]], 80, 26)
    o, _ = systems.oracle_system([('O', (0.0, 0.0, 0.0))], [[]], 80, 26)
    phi, ri, V, exact = _dpsi_closed_form_inputs(o, 0.9, 0.3, -0.6, 0.05, 1.0)
    d = napmo_amd.compute_dpsi(g, 2, phi, 0.05, -0.6, ri, V, 1.0)
    err = np.abs(d - exact).max()
    print('dpsi closed form on the GPU: max|err| %.3e' % err)
    assert err < 3e-5


def test_becke_dickson_step_matches_oracle():
    import napmo_amd
    g, _ = systems.gpu_system(ATOMS, [[]] * 3, 30, 50)
    o, _ = systems.oracle_system(ATOMS, [[]] * 3, 30, 50)
    psi, _, V = _orbital_inputs(o, 3, 9)
    T = np.array([0.4 * p * (1.0 + 0.1 * o.points[:, 1]) for p in psi])
    e = np.array([-0.9, -0.5, -0.3])
    dpsi, de, R = napmo_amd.becke_dickson_step(g, 4, psi, T, V, e, 1.0)
    rdp, rde, rR = core.becke_dickson_step(o, 4, psi, T, V, e, 1.0, evaluate=_evaluate_c)
    assert np.allclose(R, rR, rtol=0, atol=1e-15 * np.abs(rR).max())
    print('step de', de, 'ref', rde)
    assert np.allclose(de, rde, rtol=1e-12, atol=1e-15)
    for k in range(3):
        scale = np.abs(rdp[k]).max()
        err = np.abs(dpsi[k] - rdp[k]).max()
        print('step orbital %d max|dpsi| %.3e max|err| %.3e' % (k, scale, err))
        assert err < 1e-10 * scale


def test_dpsi_300_radial_points_matches_oracle():
    # the reference's numerical He grid size (300 radial points, 110 angular; lmax 8): the
    # radial system needs 84 KB of LDS per workgroup
    import napmo_amd
    atoms = [('He', (0.0, 0.0, 0.0))]
    g, _ = systems.gpu_system(atoms, [[]], 300, 110)
    o, _ = systems.oracle_system(atoms, [[]], 300, 110)
    r = np.sqrt((o.points ** 2).sum(1))
    phi = np.exp(-1.7 * r) * (1.0 + 0.2 * o.points[:, 2])
    ri = 0.05 * phi * np.cos(o.points[:, 0]) + 0.01 * np.exp(-r ** 2)
    V = -2.0 / np.maximum(r, 1e-3) * phi
    d = napmo_amd.compute_dpsi(g, 8, phi, 0.01, -0.9, ri, V, 1.0)
    ref = core.compute_dpsi(o, 8, phi, 0.01, -0.9, ri, V, 1.0, evaluate=_evaluate_c)
    scale, err = np.abs(ref).max(), np.abs(d - ref).max()
    print('dpsi 300 x 110: max|d| %.3e max|err| %.3e' % (scale, err))
    assert err < 1e-10 * scale


def test_dpsi_rejects_oversized_radial_grid():
    import napmo_amd
    g, _ = systems.gpu_system([('H', (0.0, 0.0, 0.0))], [[]], 520, 6)
    with pytest.raises(RuntimeError, match='radial grids need 3 to 512 points'):
        napmo_amd.compute_dpsi(g, 0, np.zeros(g.size), 0.0, -0.5, np.zeros(g.size),
                               np.zeros(g.size), 1.0)
