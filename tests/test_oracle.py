"""Pins the CPU oracle (oracle/core.py, oracle/napmo_oracle.c) against the reference's
own routines (golden vectors from oracle/_ref, and oracle/_ref itself when built),
the reference tests' constants, and closed forms. CPU only."""
import json
import math
import os

import numpy as np
import pytest

from oracle import core, cext
from tests import systems

GOLD = np.load(os.path.join(os.path.dirname(__file__), 'golden', 'ref_vectors.npz'))


def test_lebedev_matches_reference_generator():
    for n in [6, 14, 26, 38, 50, 74, 86, 110, 146, 170, 194, 230, 266, 302, 350, 434, 590]:
        xyz, w = core.lebedev(n)
        assert np.array_equal(xyz, GOLD['leb_%d_xyz' % n]), n
        assert np.array_equal(w, GOLD['leb_%d_w' % n]), n


@pytest.mark.parametrize('n', [50, 110, 194, 302, 590])
def test_lebedev_exactness(n):
    # sum_j aw_j Y_lm Y_l'm' = delta / (4 pi) up to the rule's degree (orthonormal Y)
    xyz, w = core.lebedev(n)
    order = core.lebedev_get_order(n)
    lmax = order // 2
    S = np.vstack([np.ones(n), core.pure_polynomials(xyz[:, 2], xyz[:, 0], xyz[:, 1], lmax)])
    lvec = np.repeat(np.arange(lmax + 1), 2 * np.arange(lmax + 1) + 1)
    Y = S * np.sqrt(2 * lvec + 1)[:, None]          # Racah -> orthonormal up to 1/sqrt(4 pi)
    G = (Y * w) @ Y.T
    assert abs(w.sum() - 1.0) < 1e-14
    assert np.allclose(G, np.eye(len(lvec)), atol=1e-12)


def test_lebedev_order_map():
    # napmo/grids/test/test_lebedev.py:10-18
    assert core.lebedev_get_order(110) == 17
    with pytest.raises(KeyError):
        core.lebedev_get_order(2)


def test_radial_transforms_match_reference():
    t = np.arange(100, dtype=float)
    cases = {'cheb': core.ChebyshevRTransform(0.5858151015155881, 100),
             'power': core.PowerRTransform(1e-3, 2e1, 100)}
    for name, rtf in cases.items():
        for which in ('radius', 'deriv', 'deriv2', 'deriv3'):
            ref = GOLD['rtf_%s_%s' % (name, which)]
            mine = getattr(rtf, which)(t)
            assert np.max(np.abs(mine - ref) / np.abs(ref)) < 1e-14, (name, which)
        r = GOLD['rtf_%s_radius' % name]
        assert np.max(np.abs(rtf.inv(r) - GOLD['rtf_%s_inv' % name])) < 1e-12


def test_he_radial_grid_golden():
    # napmo/wavefunction/test/test_kinetic.py:64-65 (He, 50 Chebyshev points)
    rdat = np.array([5.0226280914746524E-004, 2.0128708010486813E-003, 4.5433445281719814E-003,
                     8.1130928327705666E-003, 1.2749730838578131E-002, 1.8489537628088734E-002,
                     1.5873699524938556, 557.42029873820456])
    idx = [0, 1, 2, 3, 4, 5, 33, 49]
    rtf = core.ChebyshevRTransform(core.atomic_radius('He'), 50)
    r, _ = core.radial_grid(rtf)
    assert np.allclose(r[idx], rdat, rtol=1e-13)


def test_radial_gaussian_integral():
    # napmo/grids/test/test_radial.py:13-31: 4 pi int r^2 exp(-r^2/2) = (2 pi)^(3/2)
    for rtf in (core.PowerRTransform(0.0005, 1e1, 100), core.ChebyshevRTransform(core.atomic_radius('--'), 100)):
        r, rw = core.radial_grid(rtf)
        assert abs(np.sum(np.exp(-0.5 * r ** 2) * rw) - (2 * np.pi) ** 1.5) < 1e-9


def test_spline_and_ode2_match_reference():
    assert np.array_equal(core.spline_dt(GOLD['spline_y']), GOLD['spline_dt'])
    args = (GOLD['ode_b'], GOLD['ode_a'], GOLD['ode_f'], GOLD['ode_bcs'], 12)
    c1, r1 = core.build_ode2(*args)
    c2, r2 = cext.build_ode2(*args)
    assert np.array_equal(c1, GOLD['ode_coeffs']) and np.array_equal(r1, GOLD['ode_rhs'])
    assert np.array_equal(c2, GOLD['ode_coeffs']) and np.array_equal(r2, GOLD['ode_rhs'])


def test_solid_harmonics_match_reference():
    for v, ref in zip(GOLD['pure_vecs'], GOLD['pure_polys']):
        mine = core.pure_polynomials(v[2], v[0], v[1], 6)
        assert np.allclose(mine, ref[:48], rtol=1e-14, atol=1e-14)


def test_dot_multi_moments_match_reference():
    pts, f, center = GOLD['mom_points'], GOLD['mom_f'], GOLD['dec_center']
    d = pts - center
    r = np.sqrt((d * d).sum(1))
    S = np.vstack([np.ones(len(r)), core.pure_polynomials(d[:, 2] / r, d[:, 0] / r, d[:, 1] / r, 3)])
    mine = np.array([(S[:, k * 10:(k + 1) * 10] * f[k * 10:(k + 1) * 10]).sum(1) for k in range(3)])
    assert np.allclose(mine, GOLD['mom_out'], rtol=1e-13, atol=1e-14)


def _golden_splines():
    lmax = int(GOLD['dec_lmax'][0])
    rtf = core.ChebyshevRTransform(float(GOLD['dec_radii'][0]), GOLD['dec_y'].shape[1])
    return [core.Spline(GOLD['dec_y'][s], GOLD['dec_dt'][s], rtf, core.PotentialExtrapolation(int(math.isqrt(s))))
            for s in range((lmax + 1) ** 2)]


def test_eval_decomposition_matches_reference():
    sp = _golden_splines()
    mine = core.eval_decomposition(sp, GOLD['dec_center'], GOLD['dec_points'])
    c = cext.eval_decomposition(sp, GOLD['dec_center'], GOLD['dec_points'], np.zeros(len(GOLD['dec_points'])))
    ref = GOLD['dec_out']
    assert np.allclose(mine, ref, rtol=1e-13, atol=1e-13 * np.abs(ref).max())
    assert np.array_equal(c, ref)


def test_collocation_matches_reference_and_constants():
    basis = core.Basis()
    basis.add_shell((0.1, 0.2, 0.3), 0, [13.01, 1.962, 0.4446], [0.019685, 0.137977, 0.478148])
    basis.add_shell((0.1, 0.2, 0.3), 1, [0.727], [1.0])
    basis.add_shell((-0.5, 0.4, 0.0), 2, [0.8, 2.5], [0.6, 0.4])
    for phi in (core.basis_compute(basis, GOLD['gto_points']), cext.basis_compute(basis, GOLD['gto_points'])):
        assert np.allclose(phi, GOLD['gto_phi'], rtol=1e-13, atol=1e-15)
    # napmo/gto/test/test_primitive_gaussian.py:13-24, test_basis_set.py:25-29
    assert abs(core.primitive_norm(0.5, (0, 0, 0)) - 0.423777208124) < 1e-12
    assert abs(core.primitive_norm(0.5, (0, 0, 3)) - 0.309483114995) < 1e-12
    b = core.Basis()
    b.add_shell((0, 0, 0), 0, [0.5], [1.0])
    assert abs(core.basis_compute(b, np.zeros((1, 3)))[0, 0] - 0.423777208124) < 1e-12
    b = core.Basis()
    b.add_shell((0, 0, 0), 3, [0.5], [1.0])
    zzz = [k for k, lx in enumerate(core.cartesian_powers(3)) if lx == (0, 0, 3)][0]
    assert abs(core.basis_compute(b, np.ones((1, 3)))[0, zzz] - 0.069055017012) < 1e-12
    sto = systems.BASIS['STO-3G']['H'][0]
    b = core.Basis()
    b.add_shell((0, 0, 0), 0, sto['prim'], sto['cont'])
    assert abs(core.basis_compute(b, np.zeros((1, 3)))[0, 0] - 0.62824688) < 1e-8


def test_becke_partition_invariants():
    rng = np.random.default_rng(7)
    atoms = systems.random_molecule(rng, 4, 4.0, min_sep=1.2)
    grid = core.BeckeGrid(atoms, 20, 26)
    # per-point partition of unity: sum over owners of the same point = 1
    pts = grid.points[::37]
    total = np.zeros(len(pts))
    for owner in range(grid.ncenter):
        offs = np.array([0] * grid.ncenter + [len(pts)])
        offs[:owner + 1] = 0
        offs[owner + 1:] = len(pts)
        total += core.becke_weights(pts, grid.origin, grid.radii, offs)
    assert np.allclose(total, 1.0, atol=1e-13)
    # brute force from Becke (1988) at a few points
    for p in range(0, grid.size, 251):
        owner = int(np.searchsorted(grid.offsets, p, side='right') - 1)
        bf = core.becke_weights_pointwise(grid.points[p], owner, grid.origin, grid.radii)
        assert abs(bf - grid.becke_weights[p]) < 1e-12
    # C restatement equals numpy
    assert np.array_equal(cext.becke_weights(grid.points, grid.origin, grid.radii, grid.offsets),
                          grid.becke_weights)
    # permutation of centres permutes nothing but the owner blocks
    perm = [2, 0, 3, 1]
    g2 = core.BeckeGrid([atoms[k] for k in perm], 20, 26)
    for new, old in enumerate(perm):
        a = g2.becke_weights[g2.offsets[new]:g2.offsets[new + 1]]
        b = grid.becke_weights[grid.offsets[old]:grid.offsets[old + 1]]
        assert np.allclose(a, b, atol=1e-14)


def test_spherical_expansion_closed_forms():
    # napmo/grids/test/test_spherical_average.py:38-85 (1s, 1pz densities)
    rtf = core.PowerRTransform(1e-3, 2e1, 100)
    origin = np.array([0.3, -0.4, 0.2])
    ag = core.AtomicGrid(100, 110, origin, 'H', rtransform=rtf)
    d = np.sqrt(((ag.points - origin) ** 2).sum(1))
    sa = core.spherical_expansion(ag, np.exp(-2 * d) / np.pi, 4)
    assert np.abs(sa[:, 0] - np.exp(-2 * ag.r) / np.pi * np.sqrt(4 * np.pi)).max() < 1e-10
    assert np.abs(sa[:, 1:]).max() < 1e-10
    z = ag.points[:, 2] - origin[2]
    sa = core.spherical_expansion(ag, np.exp(-d) / (32.0 * np.pi) * z ** 2, 4)
    r = ag.r
    assert np.abs(sa[:, 0] - np.exp(-r) / (32 * np.pi) / 3 * r ** 2 * np.sqrt(4 * np.pi)).max() < 1e-10
    assert np.abs(sa[:, 4] - np.exp(-r) / (32 * np.pi) * 2 / 15 * r ** 2 * np.sqrt(5 * 4 * np.pi)).max() < 1e-10


def test_coulomb_self_energy_h_sto3g():
    # napmo/grids/test/test_coulomb.py:24-56 (0.77461, allclose), and the closed form
    # (ss|ss) = 2 pi^(5/2) sum c^4 / (p q sqrt(p+q)) for the one-centre 1s.
    sto = systems.BASIS['STO-3G']['H'][0]
    grid, basis = systems.oracle_system([('H', (0.0, 0.0, 0.0))], [[sto]], 100, 110)
    phi = core.basis_compute(basis, grid.points)
    rho = phi[:, 0] ** 2
    v = core.compute_coulomb(grid, rho, systems.lebedev_lmax(110))
    integral = grid.integrate(rho * v)
    assert np.allclose(integral, 0.77461)
    z = np.array(sto['prim'])
    c = np.array(sto['cont']) * np.array([core.primitive_norm(q, (0, 0, 0)) for q in z])
    c *= core.contraction_norm(sto['prim'], sto['cont'], (0, 0, 0))
    closed = sum(c[i] * c[j] * c[k] * c[l] * 2 * np.pi ** 2.5 / ((z[i] + z[j]) * (z[k] + z[l]) * np.sqrt(z[i] + z[j] + z[k] + z[l]))
                 for i in range(3) for j in range(3) for k in range(3) for l in range(3))
    assert abs(integral - closed) < 1e-5


def test_poisson_gaussian_closed_form():
    # V of a normalised s Gaussian (alpha) is erf(sqrt(alpha) r) / r. The reference's
    # finite-element scheme converges to it with the radial grid size; a dropped term,
    # a wrong sign or a wrong boundary value would leave an O(1) error instead.
    from scipy.special import erf
    alpha = 1.3
    errs = []
    for nr in (80, 150):
        ag = core.AtomicGrid(nr, 26, np.zeros(3), 'H')
        d = np.sqrt((ag.points ** 2).sum(1))
        rho = (alpha / np.pi) ** 1.5 * np.exp(-alpha * d * d)
        U = core.poisson_radial(ag, core.spherical_expansion(ag, rho, 0), 0)
        exact = erf(np.sqrt(alpha) * ag.r) / ag.r * np.sqrt(4 * np.pi)   # U_00 = sqrt(4 pi) V
        mask = ag.r < 20
        errs.append(np.max(np.abs(U[0].y[mask] - exact[mask]) / exact[mask]))
    assert errs[0] < 3e-4 and errs[1] < errs[0] / 3


def test_contraction_identities():
    rng = np.random.default_rng(3)
    atoms = systems.H2_ATOMS
    grid, basis = systems.oracle_system(atoms, systems.named_shells(atoms, 'CC-PVDZ'), 30, 50)
    phi = core.basis_compute(basis, grid.points)
    v = rng.normal(size=grid.size)
    M = core.coupling_matrix(phi, v, grid.becke_weights, grid.weights)
    Mc = cext.coupling(np.ascontiguousarray(phi.T), v, grid.becke_weights, grid.weights)
    direct = np.einsum('pi,pj,p->ij', phi, phi, v * grid.becke_weights * grid.weights)
    assert np.allclose(M, direct, rtol=1e-12, atol=1e-14) and np.allclose(Mc, direct, rtol=1e-12, atol=1e-14)
    P = rng.normal(size=(basis.nbasis, basis.nbasis))
    P = P + P.T
    rho = core.density_from_dm(phi, P)
    assert np.allclose(cext.density(phi, P), rho, rtol=1e-12, atol=1e-14)
    # trace identity: sum P o J(v) = sum_p w bw rho v
    assert np.isclose(np.sum(P * M), np.sum(grid.weights * grid.becke_weights * rho * v), rtol=1e-12)


def test_two_body_exchange_mirrors_upper_triangle():
    rng = np.random.default_rng(4)
    N, n = 40, 5
    phi, K = rng.normal(size=(N, n)), rng.normal(size=(N, n))
    J, bw, w = rng.normal(size=N), rng.uniform(size=N), rng.uniform(size=N)
    G = core.two_body_matrix(phi, J, K, -0.5, bw, w)
    E = np.einsum('pi,pj,p->ij', phi, K, bw * w)
    expect = core.coupling_matrix(phi, J, bw, w) - 0.5 * (np.triu(E) + np.triu(E, 1).T)
    assert np.allclose(G, expect)
    assert np.allclose(G, G.T)


def test_dot_multi_moments_every_mtype_matches_reference():
    # oracle/_ref vectors: two integrand vectors, segments 20/7/30, non-zero starting output
    for mtype in (1, 2, 3, 4):
        out = GOLD['mm%d_init' % mtype].copy()
        core.dot_multi_moments(GOLD['mm_data'], GOLD['mm_points'], GOLD['mm_center'], 5, mtype,
                               np.array([20, 7, 30]), out)
        assert np.array_equal(out, GOLD['mm%d_out' % mtype]), mtype
    # an empty segment keeps every later point (the reference's shift_segment walk)
    out = core.dot_multi_moments(GOLD['mm_data'][0], GOLD['mm_points'], GOLD['mm_center'], 3, 2,
                                 GOLD['mm_empty_seg'], np.zeros((4, 16)))
    assert np.array_equal(out, GOLD['mm_empty_out'])
    assert np.all(out[2:] == 0.0)
    # Cartesian order x, y, z, xx, xy, xz, yy, yz, zz at one point with unit integrand
    row = core.moment_polynomials(np.array([[2.0, 3.0, 5.0]]), 2, 1)[:, 0]
    assert np.array_equal(row, [2, 3, 5, 4, 6, 10, 9, 15, 25])
    assert np.array_equal(core.moment_polynomials(np.array([[0.0, 3.0, 4.0]]), 3, 3)[:, 0], [5, 25, 125])


def test_density_from_dm_screening_known_answer():
    # hand-computed: n = 2, lower-triangular dm, epsilon 0.01, dmmaxrow [2, 1]
    phi = np.array([[1.0, 0.001], [0.5, 1.0], [1e-3, 1e-3]])
    dm = np.array([[2.0, 99.0], [0.3, 1.0]])          # the upper entry is never read
    out = np.full(3, -7.0)
    core.density_from_dm_screened(phi, dm, 0.01, np.array([2.0, 1.0]), out)
    # point 0: eps -> 0.01 / 4, phi_1 * 1 = 1e-3 below it: rho = 2 * 1 * 1 only
    # point 1: eps -> 6.25e-4, nothing screened: 0.5 * 2 * 0.5 * ... = 0.5 + 1.3
    # point 2: bound 6e-6 below 6.25e-4: the routine returns, out[2] untouched
    assert np.array_equal(out, [2.0, 1.8, -7.0])
    # epsilon <= 0: the symmetric form diag(Phi S Phi^T), S the lower triangle mirrored
    rng = np.random.default_rng(5)
    phi = rng.normal(size=(40, 6))
    dm = rng.normal(size=(6, 6))
    S = np.tril(dm) + np.tril(dm, -1).T
    out = core.density_from_dm_screened(phi, dm, 0.0, None, np.zeros(40))
    assert np.allclose(out, np.einsum('pi,ij,pj->p', phi, S, phi), rtol=1e-13, atol=1e-13)


def test_radial_deriv2_closed_form_convergence():
    # (2/r) f' + f'' of exp(-a r^2) is (4 a^2 r^2 - 6 a) exp(-a r^2); fourth-order-like
    # convergence in the number of Chebyshev points pins the stencils and the mapping
    errs = []
    for n in (50, 100, 200):
        rtf = core.ChebyshevRTransform(1.0, n)
        r = rtf.radius(np.arange(n, dtype=np.float64))
        f = np.exp(-0.7 * r * r)
        d = core.radial_deriv2(1.0, r, f)
        ex = (4 * 0.49 * r * r - 4.2) * f
        errs.append(np.abs(d - ex)[r < 6].max())
    assert errs[2] < 3e-5 and errs[1] < errs[0] / 8 and errs[2] < errs[1] / 8


def test_kinetic_closed_form_gaussian():
    # one centre: -1/2 laplacian of exp(-a r^2) = -1/2 (4 a^2 r^2 - 6 a) exp(-a r^2); the
    # error (largest at the nucleus) falls about 16x per doubling of the radial grid
    errs = []
    for nr in (60, 120):
        grid = core.BeckeGrid([('H', (0.0, 0.0, 0.0))], nr, 50)
        r = np.sqrt((grid.points ** 2).sum(1))
        psi = np.exp(-0.8 * r * r)
        T = core.compute_kinetic(grid, psi, 2)
        ex = -0.5 * (4 * 0.64 * r * r - 4.8) * psi
        errs.append(np.max(np.abs(T - ex)))
    assert errs[1] < 2e-4 and errs[1] < errs[0] / 8


def test_nuclear_potential_matches_direct_sum():
    pts = np.array([[0.0, 0.0, 1.0], [1.0, 2.0, 2.0]])
    V = core.compute_nuclear(pts, [(1.0, (0.0, 0.0, 0.0)), (8.0, (1.0, 2.0, 0.0))])
    assert np.allclose(V, [1.0 + 8.0 / np.sqrt(1 + 4 + 1), 1.0 / 3.0 + 8.0 / 2.0], rtol=1e-15)


def test_exchange_potential_index_structure():
    # psi_numeric.py:186-215 with D = e_0 e_1^T (D[0, 1] = 1): chi_1 = psi_0, chi_0 = 0, so
    # Kpot_0 = V[psi_1 psi_0] psi_0 and Kpot_1 = V[psi_1^2] psi_0 (a transposed D would give
    # V[psi_0^2] psi_1 and V[psi_1 psi_0] psi_1 instead)
    grid = core.BeckeGrid([('H', (0.0, 0.0, 0.7)), ('H', (0.0, 0.0, -0.7))], 30, 26)
    r0 = np.sqrt(((grid.points - grid.origin[0]) ** 2).sum(1))
    r1 = np.sqrt(((grid.points - grid.origin[1]) ** 2).sum(1))
    psi = np.array([np.exp(-r0 * r0), np.exp(-0.5 * r1 * r1)])
    D = np.array([[0.0, 1.0], [0.0, 0.0]])
    K = core.compute_exchange_potential(grid, psi, D, 4)
    v10 = core.compute_coulomb(grid, psi[1] * psi[0], 4)
    v11 = core.compute_coulomb(grid, psi[1] * psi[1], 4)
    assert np.allclose(K[0], v10 * psi[0], rtol=1e-14, atol=1e-14)
    assert np.allclose(K[1], v11 * psi[0], rtol=1e-14, atol=1e-14)


def _dpsi_closed_form_inputs(grid, alpha, beta, oi, doi, mass_inv):
    # target u(r) = r^2 exp(-alpha r^2) in the l = 0 and the l = 1 (z / r) channel, with
    # V_tot = beta r^2 (its spherical average on one centre is itself). For each channel
    #   u'' + 2 u'/r + c (l(l+1)/r^2 + V - E) u = c g_l,  c = -2 / m_inv,  E = oi + doi,
    # and u'' + 2u'/r = (6 - 14 a r^2 + 4 a^2 r^4) exp(-a r^2) gives the source g_l.
    c = -2.0 / mass_inv
    p = grid.points
    r = np.sqrt((p ** 2).sum(1))
    e = np.exp(-alpha * r * r)
    lap = (6 - 14 * alpha * r * r + 4 * alpha * alpha * r ** 4) * e
    V = beta * r * r
    E = oi + doi
    g0 = (lap + c * (V - E) * r * r * e) / c
    g1 = (lap + c * (2.0 + (V - E) * r * r) * e) / c
    src = g0 + g1 * p[:, 2] / r
    phi = np.exp(-0.3 * r)             # any orbital: ri = phi doi - src
    ri = phi * doi - src
    exact = r * r * e * (1.0 + p[:, 2] / r)
    return phi, ri, V, exact


@pytest.mark.parametrize('mass_inv', [1.0, 0.5])
def test_dpsi_closed_form_convergence(mass_inv):
    # ndpsi.py:15-89 on one centre against the exact solution of its radial equations
    # (doubled centrifugal term, -2/m_inv scaling, V_sph - oi - doi, zero boundary values):
    # a dropped factor or sign in any of them moves the l = 0 or l = 1 channel by O(1)
    errs = []
    for nr in (40, 80):
        grid = core.BeckeGrid([('O', (0.0, 0.0, 0.0))], nr, 26)
        phi, ri, V, exact = _dpsi_closed_form_inputs(grid, 0.9, 0.3, -0.6, 0.05, mass_inv)
        d = core.compute_dpsi(grid, 2, phi, 0.05, -0.6, ri, V, mass_inv)
        errs.append(np.max(np.abs(d - exact)))
    # fourth order in the radial spacing (1.9e-4 -> 1.2e-5 for max|u| = 0.82 at m_inv = 1)
    print('dpsi closed form errors', errs)
    assert errs[1] < 3e-5 and errs[1] < errs[0] / 8


def test_becke_dickson_step_residual_and_energy():
    # psi_optimization.py:175-203: R = T + V - e psi, de = integrate(R psi); dpsi uses them
    grid = core.BeckeGrid([('H', (0.0, 0.0, 0.7)), ('H', (0.0, 0.0, -0.7))], 20, 26)
    r0 = np.sqrt(((grid.points - grid.origin[0]) ** 2).sum(1))
    r1 = np.sqrt(((grid.points - grid.origin[1]) ** 2).sum(1))
    psi = np.array([np.exp(-r0) + np.exp(-r1)])
    T = np.array([0.3 * psi[0]])
    V = np.array([-(1 / r0 + 1 / r1) * psi[0]])
    e = np.array([-0.55])
    dpsi, de, R = core.becke_dickson_step(grid, 2, psi, T, V, e, 1.0)
    assert np.array_equal(R[0], T[0] + V[0] - e[0] * psi[0])
    assert de[0] == grid.integrate(R[0] * psi[0])
    ref = core.compute_dpsi(grid, 2, psi[0], de[0], e[0], R[0], V[0], 1.0)
    assert np.array_equal(dpsi[0], ref)


def test_common_index_known_answer():
    # multi_grid.cpp:37-86: rows (i in the larger grid, j), i-major then j; -0.0 == 0.0;
    # a duplicated point of the smaller grid is listed for each of its copies
    A = np.array([[0.0, 0.0, 0.0], [1.0, 2.0, -0.0], [5.0, 5.0, 5.0], [0.5, 0.25, 3.0],
                  [9.0, 9.0, 9.0]])
    B = np.array([[0.5, 0.25, 3.0], [7.0, 7.0, 7.0], [1.0, 2.0, 0.0], [0.5, 0.25, 3.0]])
    idx = core.common_index(A, B)
    assert idx.tolist() == [[1, 2], [3, 0], [3, 3]]
    # a smaller first grid is swapped: column 0 then indexes B
    idx = core.common_index(A[:3], B)
    assert idx.tolist() == [[2, 1]]
    assert core.common_index(A[2:3], B).shape == (0, 2)


def test_epc17_2_closed_form_and_derivatives():
    # the published form eps = -rN / (2.35 - 2.4 sqrt(rE rN) + 6.6 rE rN) at rE = rN = 1:
    # 2.35 - 2.4 + 6.6 = 6.55
    e, _, _ = core.epc17_2(np.array([1.0]), np.array([1.0]))
    assert e[0] == -1.0 / 6.55
    # both potential forms are the derivatives of rE eps (central differences)
    rng = np.random.default_rng(2)
    rE = rng.uniform(0.01, 3.0, 50)
    rN = rng.uniform(0.01, 3.0, 50)
    for small in (False, True):
        xE, xN = rE.copy(), rN.copy()
        if small:       # one zero density anywhere switches the whole array's form
            xE = np.append(xE, 0.0)
            xN = np.append(xN, 1.0)
        _, pA, pB = core.epc17_2(xE, xN)
        h = 1e-6
        f = lambda a, b: a * core.epc17_2(a, b)[0]  # noqa: E731
        with np.errstate(invalid='ignore'):   # the appended zero density goes negative
            dA = (f(xE + h, xN) - f(xE - h, xN)) / (2 * h)
            dB = (f(xE, xN + h) - f(xE, xN - h)) / (2 * h)
        n = len(rE)
        assert np.allclose(pA[:n], dA[:n], rtol=1e-7, atol=1e-9)
        assert np.allclose(pB[:n], dB[:n], rtol=1e-7, atol=1e-9)


def test_lowdin2_grid_file(tmp_path):
    from scipy.io import FortranFile
    rng = np.random.default_rng(4)
    pw = rng.normal(size=(7, 4))
    path = str(tmp_path / 'grid.dat')
    with FortranFile(path, 'w') as f:
        f.write_record(np.array([pw.size], dtype=np.int64))
        f.write_record(np.asfortranarray(pw).ravel(order='F'))
    out = core.read_lowdin2_grid(path)
    assert np.array_equal(out, pw)


def test_c_2species_grid_scatter_and_energy():
    index = np.array([[2, 0], [0, 1]])
    rho_a = np.array([0.5, 0.1, 2.0])
    rho_b = np.array([1.5, 0.7])
    bw, w = np.array([0.9, 0.4]), np.array([2.0, 3.0])
    va, vb = np.zeros(3), np.zeros(2)
    E = core.c_2species_grid(index, rho_a, rho_b, bw, w, va, vb)
    ene, pA, pB = core.epc17_2(np.array([2.0, 0.5]), np.array([1.5, 0.7]))
    assert np.isclose(E, 0.9 * 2.0 * ene[0] * 2.0 + 0.4 * 3.0 * ene[1] * 0.5, rtol=1e-15)
    assert va.tolist() == [pA[1], 0.0, pA[0]] and vb.tolist() == [pB[0], pB[1]]
