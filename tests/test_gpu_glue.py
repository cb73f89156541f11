"""Cross-species glue on the GPU: Lowdin2 point grids, the hashed MultiGrid common-point
index and the epc17-2 correlation on common points, against the oracle."""
import numpy as np
import pytest

from oracle import core
from tests import systems

pytestmark = pytest.mark.gpu

E_ATOMS = [('O', (0.0, 0.0, 0.0)), ('H', (1.1, 0.9, 0.0)), ('H', (-1.2, 0.7, 0.2))]


def _species_grids():
    """Electrons on three atoms (20 x 26) and a quantum proton on the first H's atomic grid:
    its points coincide exactly with that atom's points of the electronic grid."""
    import napmo_amd
    ge = napmo_amd.BeckeGrid(napmo_amd.make_species(E_ATOMS, symbol='e-'), 20, 26)
    gp = napmo_amd.BeckeGrid(napmo_amd.make_species([E_ATOMS[1]], symbol='H+'), 20, 26)
    return ge, gp


def test_multigrid_common_points_match_oracle():
    import napmo_amd
    ge, gp = _species_grids()
    mg = napmo_amd.MultiGrid(2)
    mg.add_grid(ge)
    mg.add_grid(gp)
    assert mg.ngrids == 2 and mg.nspecies == 2
    from napmo_amd.system import cext
    assert cext.MultiGrid_get_common_index_size(mg._this, 0, 1) == -1
    idx = np.array(mg.get_common_points('e-', 'H+'))
    ref = core.common_index(ge.points, gp.points)
    print('common points', idx.shape, 'of', ge.size, gp.size)
    assert idx.shape == (gp.size, 2) and np.array_equal(idx, ref)
    # the proton grid is the electron grid's second atom, point for point
    assert np.array_equal(idx[:, 1], np.arange(gp.size))
    assert np.array_equal(idx[:, 0], np.arange(gp.size) + 20 * 26)
    # (b, a) gives the same rows (column 0 is always the larger grid)
    assert np.array_equal(np.array(mg.get_common_points('H+', 'e-')), ref)
    assert np.array_equal(mg.get_common_points('e-', 'e-'), np.arange(ge.size))


def test_lowdin2_grid_from_file(tmp_path):
    import napmo_amd
    from napmo_amd.system import cext
    from scipy.io import FortranFile
    ge, _ = _species_grids()
    pw = np.concatenate([ge.points, ge.becke_weights[:, None]], axis=1)
    path = str(tmp_path / 'lowdin2.grid')
    with FortranFile(path, 'w') as f:
        f.write_record(np.array([pw.size], dtype=np.int64))
        f.write_record(np.asfortranarray(pw).ravel(order='F'))
    g = napmo_amd.BeckeGrid(napmo_amd.make_species(E_ATOMS, symbol='e-'), file=path)
    assert g.size == ge.size and g.ncenter == 3
    assert np.array_equal(g.points, ge.points)
    assert np.array_equal(g.becke_weights, ge.becke_weights)
    assert np.array_equal(g.weights, np.ones(g.size))
    f = np.cos(g.points[:, 0])
    assert abs(g.integrate(f) - np.sum(f * ge.becke_weights)) < 1e-12 * np.abs(f).sum()
    # no atomic grids: the Poisson path reports it instead of running
    out = np.zeros(g.size)
    assert cext.napmo_compute_coulomb(g._this, np.ones(g.size), 2, out) != 0
    assert 'points only' in cext.lib.napmo_last_error().decode()


def test_epc17_2_matches_oracle():
    import napmo_amd
    rng = np.random.default_rng(6)
    rE = 10.0 ** rng.uniform(-8, 1, 4096)
    rN = 10.0 ** rng.uniform(-8, 1, 4096)
    for small in (False, True):
        a, b = rE.copy(), rN.copy()
        if small:
            a[17] = 0.0
        out = napmo_amd.epc17_2(a, b)
        ref = core.epc17_2(a, b)
        # error relative to the size of the terms each potential is a difference of (the
        # simplified form's pow(r, 1.5) may differ from the host's in the last bit, and
        # its two terms can nearly cancel)
        den = 2.35 - 2.4 * np.sqrt(a * b) + 6.6 * a * b
        scales = (np.abs(ref[0]),
                  (2.4 * np.sqrt(a) * b ** 1.5 + 4.7 * b) / den ** 2 / 2 + np.abs(ref[1]),
                  (2.4 * np.sqrt(b) * a ** 1.5 + 4.7 * a) / den ** 2 / 2 + np.abs(ref[2]))
        for x, y, sc in zip(out, ref, scales):
            err = np.max(np.abs(x - y) / np.maximum(sc, 1e-300))
            print('epc17-2 small=%s max rel err %.2e' % (small, err))
            assert err < 4e-15


def test_c_2species_grid_matches_oracle():
    import napmo_amd
    from napmo_amd.functional import c_2species_grid
    ge, gp = _species_grids()
    mg = napmo_amd.MultiGrid(2)
    mg.add_grid(ge)
    mg.add_grid(gp)
    idx = np.array(mg.get_common_points('e-', 'H+'))
    pe, pp = ge.points, gp.points
    rho_e = sum(np.exp(-((pe - np.array(c)) ** 2).sum(1)) for _, c in E_ATOMS)
    rho_p = 30.0 * np.exp(-8.0 * ((pp - np.array(E_ATOMS[1][1])) ** 2).sum(1))
    va, vb = np.full(ge.size, 0.25), np.full(gp.size, -0.5)
    E = c_2species_grid(ge, gp, idx, rho_e, rho_p, va, vb)
    ra, rb = np.full(ge.size, 0.25), np.full(gp.size, -0.5)
    Eref = core.c_2species_grid(idx, rho_e, rho_p, gp.becke_weights, gp.weights, ra, rb)
    print('epc17-2 energy', E, 'ref', Eref)
    assert abs(E - Eref) < 1e-13 * abs(Eref)
    assert np.allclose(va, ra, rtol=1e-14, atol=0) and np.allclose(vb, rb, rtol=1e-14, atol=0)
    with pytest.raises(RuntimeError, match='out of range'):
        bad = idx.copy()
        bad[0, 1] = gp.size
        c_2species_grid(ge, gp, bad, rho_e, rho_p, va, vb)
