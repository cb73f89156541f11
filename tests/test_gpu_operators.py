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
    lmax = 6
    L = (lmax + 1) ** 2
    # the radial half: spline values and slopes per (atom, l, m) against the oracle's
    from napmo_amd.system import cext
    nrad = [ag.r.size for ag in o.atgrids]
    Uy, Ud, Rl = np.zeros(sum(nrad) * L), np.zeros(sum(nrad) * L), np.zeros(sum(nrad) * L)
    cext.check(cext.napmo_kinetic_expand(g._this, np.ascontiguousarray(psi), lmax, Uy, Ud, Rl.ctypes.data))
    off = 0
    worst_l = np.zeros(lmax + 1)     # GPU y against the oracle chain
    worst_rho = np.zeros(lmax + 1)   # GPU projection against the oracle's
    worst_rad = np.zeros(lmax + 1)   # GPU y against the oracle deriv2 of the GPU projection
    for k, ag in enumerate(o.atgrids):
        f = psi[o.offsets[k]:o.offsets[k + 1]] * o.becke_weights[o.offsets[k]:o.offsets[k + 1]]
        sph = core.spherical_expansion(ag, f, lmax)
        r = ag.r
        yg = Uy[off * L:(off + r.size) * L].reshape(r.size, L)
        rg = Rl[off * L:(off + r.size) * L].reshape(r.size, L)
        for lm in range(L):
            l = int(np.sqrt(lm))
            y = core.radial_deriv2(ag.rtf.radii, r, sph[:, lm]) - l * (l + 1) / (r * r) * sph[:, lm]
            yr = core.radial_deriv2(ag.rtf.radii, r, rg[:, lm]) - l * (l + 1) / (r * r) * rg[:, lm]
            sc = max(np.abs(y).max(), 1e-300)
            worst_l[l] = max(worst_l[l], np.abs(yg[:, lm] - y).max() / sc)
            worst_rad[l] = max(worst_rad[l], np.abs(yg[:, lm] - yr).max() / sc)
            # relative to the size of rho_00: the l >= 1 components are small sums that cancel
            worst_rho[l] = max(worst_rho[l], np.abs(rg[:, lm] - sph[:, lm]).max() / np.abs(sph[:, 0]).max())
        off += r.size
    print('kinetic expansion worst rel per l', worst_l)
    print('kinetic projection worst rel per l', worst_rho)
    print('kinetic radial step worst rel per l (same projection)', worst_rad)
    T = napmo_amd.compute_kinetic(g, psi, lmax)
    ref = core.compute_kinetic(o, psi, lmax)
    i = int(np.argmax(np.abs(T - ref)))
    print('kinetic max|dT| %.3e max|T| %.3e at point %d (distances %s) T %.12e ref %.12e' %
          (np.abs(T - ref).max(), np.abs(ref).max(), i,
           [round(float(np.linalg.norm(o.points[i] - np.array(c))), 6) for _, c in atoms], T[i], ref[i]))
    # The projection agrees to rounding (relative to rho_00; the small l >= 2 components,
    # whose angular sums cancel, carry 1e-12..1e-10 of their own size) and the radial step
    # applied to the same projection agrees to 1e-9. The finite-difference Laplacian of the reference (deriv2 - l(l+1)/r^2 near
    # r -> 0) amplifies the projection's rounding about 1e5-fold, so the full operator
    # agrees to ~1e-6 of max|T|, worst at points within the first radial node of a nucleus.
    assert np.max(worst_rho) < 1e-12
    assert np.max(worst_rad) < 1e-9
    assert np.max(np.abs(T - ref)) < 1e-5 * np.abs(ref).max()
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


def test_exchange_potential_matches_oracle():
    import napmo_amd
    # H2 cc-pVDZ on 2 x 30 x 50 points, lmax 6: 4 functions, 10 pair potentials
    atoms, shells = systems.H2_ATOMS, systems.named_shells(systems.H2_ATOMS, 'CC-PVDZ')
    g, b = systems.gpu_system(atoms, shells, 30, 50)
    o, ob = systems.oracle_system(atoms, shells, 30, 50)
    from oracle import cext as ocx
    psi = np.ascontiguousarray(ocx.basis_compute(ob, o.points)[:, :4].T)
    D = np.random.default_rng(8).normal(size=(4, 4))
    K = napmo_amd.compute_exchange_potential(g, psi, D, 6)
    ref = core.compute_exchange_potential(o, psi, D, 6)
    print('exchange max|dK| %.3e max|K| %.3e' % (np.abs(K - ref).max(), np.abs(ref).max()))
    assert np.max(np.abs(K - ref)) < 1e-11 * np.abs(ref).max()


def test_device_entry_points_match_host_ones():
    # napmo_compute_{exchange_potential,kinetic}_device on resident arrays give the host
    # entry points' results bit for bit
    import napmo_amd
    from napmo_amd.device import DeviceArray
    from napmo_amd.system import cext
    atoms = [('O', (0.0, 0.0, 0.0)), ('H', (1.1, 0.9, 0.0))]
    g, _ = systems.gpu_system(atoms, [[]] * 2, 24, 50)
    p = g.points
    psi = np.array([np.exp(-0.7 * ((p - np.array(c)) ** 2).sum(1)) * (1 + 0.2 * k * p[:, 0])
                    for k, (_, c) in enumerate(atoms * 2)][:3])
    D = np.array([[1.0, 0.2, 0.0], [0.2, 0.5, 0.1], [0.0, 0.1, 0.8]])
    K = napmo_amd.compute_exchange_potential(g, psi, D, 4)
    T = napmo_amd.compute_kinetic(g, psi[1], 4)
    dpsi, dD, dK, dT = DeviceArray.from_host(psi), DeviceArray.from_host(D), DeviceArray(psi.shape), DeviceArray((g.size,))
    cext.check(cext.napmo_compute_exchange_potential_device(g._this, 3, dpsi.ptr, dD.ptr, 4, dK.ptr))
    cext.check(cext.napmo_compute_kinetic_device(g._this, dpsi.ptr + 8 * g.size, 4, dT.ptr))
    assert np.array_equal(dK.to_host(), K)
    assert np.array_equal(dT.to_host(), T)
