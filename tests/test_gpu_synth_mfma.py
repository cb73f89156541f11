"""The many-density synthesis on FP64 MFMA (used by the exchange potential and the orbital
corrections) against the one-density bucketed synthesis and the oracle: a full tile of 16
densities plus a partial tile, interval and extrapolation buckets, lmax 4 and 12, and
run-to-run reproducibility. The grid knob NAPMO_TUNE_SYNTH_MFMA_MIN (9) chooses the path:
1 sends every density through the MFMA kernel, 17 none."""
import numpy as np
import pytest

from oracle import core
from tests import systems
from tests.test_gpu_dpsi import ATOMS, _evaluate_c, _orbital_inputs

pytestmark = pytest.mark.gpu

MFMA_MIN = 9


def _both(grid, fn):
    """fn() with every density on MFMA, then with every density one at a time."""
    grid.set_tuning(MFMA_MIN, 1)
    assert grid.get_tuning(MFMA_MIN) == 1
    a = fn()
    grid.set_tuning(MFMA_MIN, 17)
    b = fn()
    grid.set_tuning(MFMA_MIN, 0)
    assert grid.get_tuning(MFMA_MIN) == 10
    return a, b


def test_dpsi_twenty_orbitals_mfma_matches_one_at_a_time_and_oracle():
    # three centres, 30 x 50 points, lmax 4; 20 orbitals = a tile of 16 and a tile of 4
    import napmo_amd
    g, _ = systems.gpu_system(ATOMS, [[]] * 3, 30, 50)
    o, _ = systems.oracle_system(ATOMS, [[]] * 3, 30, 50)
    norb = 20
    phi, ri, V = _orbital_inputs(o, norb, 6)
    doi, oi = np.linspace(0.03, -0.02, norb), np.linspace(-0.7, -0.3, norb)
    run = lambda: napmo_amd.compute_dpsi(g, 4, phi, doi, oi, ri, V, 1.0)  # noqa: E731
    a, b = _both(g, run)
    scale = np.abs(b).max()
    err = np.abs(a - b).max()
    print('dpsi 20 orbitals: max|MFMA - one at a time| %.3e (max|dpsi| %.3e)' % (err, scale))
    assert err < 1e-13 * scale
    for k in (0, 15, 16, 19):
        ref = core.compute_dpsi(o, 4, phi[k], doi[k], oi[k], ri[k], V[k], 1.0, evaluate=_evaluate_c)
        e = np.abs(a[k] - ref).max()
        print('  orbital %d vs oracle %.3e' % (k, e))
        assert e < 1e-10 * np.abs(ref).max()
    g.set_tuning(MFMA_MIN, 1)
    assert np.array_equal(run(), a)
    g.set_tuning(MFMA_MIN, 0)


def test_exchange_lmax12_mfma_matches_one_at_a_time_and_oracle():
    # three centres on 20 x 302 points, lmax 12: 6 functions -> 21 pair densities
    import napmo_amd
    from oracle import cext as ocx
    shells = systems.named_shells(ATOMS, 'STO-3G')
    g, _ = systems.gpu_system(ATOMS, shells, 20, 302)
    o, ob = systems.oracle_system(ATOMS, shells, 20, 302)
    psi = np.ascontiguousarray(ocx.basis_compute(ob, o.points)[:, :6].T)
    D = np.random.default_rng(9).normal(size=(6, 6))
    a, b = _both(g, lambda: napmo_amd.compute_exchange_potential(g, psi, D, 12))
    scale = np.abs(b).max()
    err = np.abs(a - b).max()
    print('exchange lmax 12: max|MFMA - one at a time| %.3e (max|K| %.3e)' % (err, scale))
    assert err < 1e-13 * scale
    ref = core.compute_exchange_potential(o, psi, D, 12)
    e = np.abs(a - ref).max()
    print('  vs oracle %.3e' % e)
    assert e < 1e-11 * np.abs(ref).max()


def test_extrapolation_buckets_mfma_matches_one_at_a_time():
    # radial grids r_i = 0.01 (i + 1)^p from 0.01 to 1.5 bohr (10 nodes): the second centre
    # sits 3e-3 bohr beyond the first centre's sixth node on the x axis, so its sixth shell's
    # -x point lies inside the first centre's first node (left extrapolation, d^l); outer
    # shells reach beyond the other centres' last nodes (right, d^-(l+1)). Both kinds of
    # bucket are checked to be populated.
    import napmo_amd
    from napmo_amd.grids.radial_transform import PowerRadialTransform
    rt = PowerRadialTransform(0.01, 1.5, 10)
    r5 = rt.rmin * 6.0 ** rt.power
    atoms = [('O', (0.0, 0.0, 0.0)), ('H', (r5 + 3e-3, 0.0, 0.0)), ('N', (0.4, 1.9, 0.3))]
    grid = napmo_amd.BeckeGrid(napmo_amd.make_species(atoms), 10, 110, rtransform=rt)
    pts = grid.points
    left = right = 0
    for _, c in atoms:
        d = np.sqrt(((pts - np.array(c)) ** 2).sum(1))
        left += int((d < rt.rmin).sum())
        right += int((d > rt.rmax).sum())
    print('pairs left of the first node %d, right of the last %d' % (left, right))
    assert left > 0 and right > 0
    rng = np.random.default_rng(11)
    psi = np.ascontiguousarray(np.exp(-rng.uniform(0.3, 1.2, (5, 1)) * (pts ** 2).sum(1)[None, :]))
    D = rng.normal(size=(5, 5))
    a, b = _both(grid, lambda: napmo_amd.compute_exchange_potential(grid, psi, D, 6))
    scale = np.abs(b).max()
    err = np.abs(a - b).max()
    print('extrapolation grid: max|MFMA - one at a time| %.3e (max|K| %.3e)' % (err, scale))
    assert err < 1e-13 * scale


@pytest.mark.parametrize('norb', [1, 15, 17])
def test_tile_boundaries(norb):
    # a lone density, a tile short of one, and a full tile plus one: every threshold gives
    # the one-at-a-time result to rounding, and the default threshold's split is reproducible
    import napmo_amd
    g, _ = systems.gpu_system(ATOMS, [[]] * 3, 20, 50)
    o, _ = systems.oracle_system(ATOMS, [[]] * 3, 20, 50)
    phi, ri, V = _orbital_inputs(o, norb, 8)
    doi, oi = np.linspace(0.02, -0.01, norb), np.linspace(-0.6, -0.3, norb)
    run = lambda: napmo_amd.compute_dpsi(g, 4, phi, doi, oi, ri, V, 1.0)  # noqa: E731
    a, b = _both(g, run)
    assert np.abs(a - b).max() < 1e-13 * np.abs(b).max()
    d0 = run()
    assert np.array_equal(d0, run())
    print('norb %d: max|MFMA - one at a time| %.2e' % (norb, np.abs(a - b).max()))
