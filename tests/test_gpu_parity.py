"""GPU parity: every hot-path kernel through the C-ABI against the CPU oracle on the same
seeded inputs. FP64 mode: tolerances are rounding-level (stated per test)."""
import math
import os

import numpy as np
import pytest

from oracle import cext as ocx
from oracle import core
from tests import systems

pytestmark = pytest.mark.gpu

GOLD = np.load(os.path.join(systems.ROOT, 'tests', 'golden', 'ref_vectors.npz'))


def _mol(seed, natom=4, box=4.0):
    return systems.random_molecule(np.random.default_rng(seed), natom, box, min_sep=1.2)


def test_becke_weights_match_oracle():
    atoms = _mol(11)
    g, _ = systems.gpu_system(atoms, [[]] * len(atoms), 30, 50)
    o = core.BeckeGrid(atoms, 30, 50)
    print('becke N', g.size)
    assert np.array_equal(g.points, o.points)
    # FP64 rounding, amplified by the cancellation in 1 - s (s -> 1) of the reference formula
    assert np.max(np.abs(g.becke_weights - o.becke_weights)) < 1e-12


def test_becke_weights_uneven_grids():
    atoms = _mol(12, natom=3)
    sizes = [(20, 26), (35, 110), (15, 50)]
    g, _ = systems.gpu_system(atoms, [[]] * 3, None, None, sizes=sizes)
    o = core.BeckeGrid(atoms, None, None, sizes=sizes)
    assert g.size == o.size == sum(a * b for a, b in sizes)
    assert np.max(np.abs(g.becke_weights - o.becke_weights)) < 1e-12
    # partition of unity on shared points: sum over owners at one physical point
    pts = o.points[::53]
    tot = sum(core.becke_weights(pts, o.origin, o.radii, [0] * (k + 1) + [len(pts)] * (3 - k)) for k in range(3))
    assert np.allclose(tot, 1.0, atol=1e-13)


def test_collocation_matches_oracle():
    atoms = _mol(13)
    rng = np.random.default_rng(13)
    shells = systems.random_shells(rng, len(atoms), 6)
    g, b = systems.gpu_system(atoms, shells, 12, 26)
    o, ob = systems.oracle_system(atoms, shells, 12, 26)
    phi = b.compute(g.points)
    ref = ocx.basis_compute(ob, o.points)
    print('collocation', phi.shape)
    scale = np.abs(ref).max(axis=0)
    assert np.max(np.abs(phi - ref) / scale) < 1e-13


def test_spherical_expansion_matches_oracle_and_closed_forms():
    import napmo_amd
    rtf = napmo_amd.PowerRadialTransform(1e-3, 2e1, 100)
    origin = np.array([0.3, -0.4, 0.2])
    ag = napmo_amd.AtomicGrid(100, 110, origin, 'H', rtransform=rtf)
    d = np.sqrt(((ag.points - origin) ** 2).sum(1))
    r = ag.radial_grid.points
    sa = ag.spherical_expansion(4, np.exp(-2 * d) / np.pi)
    assert np.abs(sa[:, 0] - np.exp(-2 * r) / np.pi * np.sqrt(4 * np.pi)).max() < 1e-10
    assert np.abs(sa[:, 1:]).max() < 1e-10
    z = ag.points[:, 2] - origin[2]
    f = np.exp(-d) / (32.0 * np.pi) * z ** 2
    sa = ag.spherical_expansion(6, f)
    assert np.abs(sa[:, 4] - np.exp(-r) / (32 * np.pi) * 2 / 15 * r ** 2 * np.sqrt(20 * np.pi)).max() < 1e-10
    og = core.AtomicGrid(100, 110, origin, 'H', rtransform=core.PowerRTransform(1e-3, 2e1, 100))
    ref = core.spherical_expansion(og, f, 6)
    assert np.max(np.abs(sa - ref)) < 1e-13 * np.abs(ref).max()


def _h2_system(nrad=100, nang=110):
    atoms = systems.H2_ATOMS
    shells = systems.named_shells(atoms, 'CC-PVDZ')
    return atoms, shells, nrad, nang


def test_poisson_solver_matches_oracle():
    import napmo_amd
    atoms, shells, nr, na = _h2_system(60, 110)
    g, b = systems.gpu_system(atoms, shells, nr, na)
    o, ob = systems.oracle_system(atoms, shells, nr, na)
    phi = ocx.basis_compute(ob, o.points)
    rho = (phi[:, 0] + 0.3 * phi[:, 4] + 0.2 * phi[:, 7]) ** 2
    lmax = 10
    U = napmo_amd.poisson_solver(g, rho, lmax)
    Uo = core.poisson_solver(o, rho, lmax)
    worst = 0.0
    for a in range(2):
        for lm in range(len(Uo[a])):
            ys, yo = U[a][lm].y, Uo[a][lm].y
            scale = max(np.abs(Uo[a][0].y).max(), 1e-30)
            worst = max(worst, np.max(np.abs(ys - yo)) / scale, np.max(np.abs(U[a][lm].dt - Uo[a][lm].dt)) / scale)
    print('poisson worst rel', worst)
    assert worst < 1e-11


def test_compute_coulomb_matches_oracle_and_reference_constant():
    import napmo_amd
    sto = systems.BASIS['STO-3G']['H'][0]
    atoms = [('H', (0.0, 0.0, 0.0))]
    g, b = systems.gpu_system(atoms, [[sto]], 100, 110)
    rho = b.compute(g.points)[:, 0] ** 2
    v = napmo_amd.compute_coulomb(g, rho, 8)
    integral = g.integrate(rho * v)
    print('coulomb integral', integral)
    assert np.allclose(integral, 0.77461)          # napmo/grids/test/test_coulomb.py:56
    o, ob = systems.oracle_system(atoms, [[sto]], 100, 110)
    vo = core.compute_coulomb(o, rho, 8)
    assert np.max(np.abs(v - vo)) < 1e-11 * np.abs(vo).max()
    # the reference's own composition (poisson_solver + evaluate_decomposition) agrees
    U = napmo_amd.poisson_solver(g, rho, 8)
    v2 = np.zeros(g.size)
    g.evaluate_decomposition(0, U[0], v2)
    assert np.max(np.abs(v2 - vo)) < 1e-11 * np.abs(vo).max()


def test_eval_decomposition_matches_reference_vectors():
    import ctypes
    import napmo_amd
    from napmo_amd.system import cext
    lmax = int(GOLD['dec_lmax'][0])
    nrad = GOLD['dec_y'].shape[1]
    rtf = napmo_amd.ChebyshevRadialTransform(float(GOLD['dec_radii'][0]), nrad)
    sp = [napmo_amd.CubicSpline.from_dt(GOLD['dec_y'][s], GOLD['dec_dt'][s], rtf,
                                        napmo_amd.PotentialExtrapolation(int(math.isqrt(s))))
          for s in range((lmax + 1) ** 2)]
    pts = np.ascontiguousarray(GOLD['dec_points'])
    out = np.zeros(len(pts))
    cext.eval_decomposition_grid(np.array([s._this for s in sp], dtype=ctypes.c_void_p),
                                 np.ascontiguousarray(GOLD['dec_center']), out, pts, napmo_amd.Cell()._this,
                                 len(sp), len(pts))
    ref = GOLD['dec_out']
    assert np.max(np.abs(out - ref)) < 1e-12 * np.abs(ref).max()


def test_decomposition_resynthesis():
    # napmo/grids/test/test_decomposition.py:20-64 (one O atom, 6-31G; density expanded,
    # splined with the default cusp extrapolation and re-synthesised on the grid)
    import napmo_amd
    atoms = [('O', (0.0, 0.0, -1.0))]
    shells = [systems.BASIS['6-31G']['O']]
    for nrad in (110, 170):
        g, b = systems.gpu_system(atoms, shells, nrad, 110)
        lmax = systems.lebedev_lmax(110)
        rho = b.compute(g.points).sum(axis=1) * g.becke_weights
        p_lm = g.atgrids[0].spherical_expansion(lmax, rho)
        rtf = g.atgrids[0].radial_grid.rtransform
        splines = [napmo_amd.CubicSpline(np.array(p_lm[:, k]), rtransform=rtf) for k in range(p_lm.shape[1])]
        rho_n = np.zeros(g.size)
        g.evaluate_decomposition(0, splines, rho_n)
        assert np.allclose(rho_n, rho)


def test_contractions_match_oracle():
    import napmo_amd
    from napmo_amd.wavefunction import grid_ops
    atoms, shells, nr, na = _h2_system(40, 50)
    g, b = systems.gpu_system(atoms, shells, nr, na)
    o, ob = systems.oracle_system(atoms, shells, nr, na)
    rng = np.random.default_rng(21)
    phi = ocx.basis_compute(ob, o.points)                   # [N, n]
    n = phi.shape[1]
    J = rng.normal(size=g.size)
    K = rng.normal(size=(g.size, n))
    psi = napmo_amd.PSIB(n, eta=2.0, kappa=-1.0)
    res = grid_ops.compute_coupling(psi, g, np.ascontiguousarray(phi.T), J)
    ref = core.coupling_matrix(phi, J, o.becke_weights, o.weights)
    assert np.max(np.abs(res - ref)) < 1e-12 * np.abs(ref).max()
    G = grid_ops.compute_2body_matrix(psi, g, np.ascontiguousarray(phi.T), J, np.ascontiguousarray(K.T))
    Gref = core.two_body_matrix(phi, J, K, psi.x_factor, o.becke_weights, o.weights)
    assert np.max(np.abs(G - Gref)) < 1e-12 * np.abs(Gref).max()
    XC = grid_ops.compute_xc_matrix(psi, g, np.ascontiguousarray(phi.T), J)
    assert np.max(np.abs(XC - ref)) < 1e-12 * np.abs(ref).max()


def _h2_fixture():
    atoms, shells, nr, na = _h2_system(100, 110)
    o, ob = systems.oracle_system(atoms, shells, nr, na)
    phi = ocx.basis_compute(ob, o.points)
    S = core.coupling_matrix(phi, np.ones(o.size), o.becke_weights, o.weights)
    P = systems.density_matrix(np.random.default_rng(1), S, 1)
    return atoms, shells, nr, na, o, phi, P


def test_grid_fock_fp64_h2_energy_parity():
    # H2 (R = 1.4 bohr), 2 x 100 x 110 Chebyshev/Lebedev points, cc-pVDZ (n = 10), lmax 10
    import napmo_amd
    atoms, shells, nr, na, o, phi, P = _h2_fixture()
    g, b = systems.gpu_system(atoms, shells, nr, na)
    fock = napmo_amd.GridFock(g, b, 10, 'fp64')
    F, E, Egrid, rho, v = fock.build(P, want_grid=True)
    rho_o, v_o, F_o, E_o = core.grid_fock_build(o, phi, P, 10)
    print('H2 E_gpu %.15f E_oracle %.15f dE %.3e' % (E, E_o, E - E_o))
    assert abs(E - E_o) < 1e-8                       # Eh
    assert np.max(np.abs(F - F_o)) < 1e-8            # Eh
    assert abs(Egrid - E) < 1e-10
    assert np.max(np.abs(rho - rho_o)) < 1e-13 * np.abs(rho_o).max()
    assert np.max(np.abs(v - v_o)) < 1e-11 * np.abs(v_o).max()
    assert abs(g.integrate(rho) - 2.0) < 1e-12        # int rho = tr(P S) = 2 electrons
    # repeated builds are deterministic
    F2, E2, _ = fock.build(P)
    assert np.array_equal(F, F2) and E == E2


def test_grid_fock_linearity_and_symmetry():
    import napmo_amd
    atoms, shells, nr, na, o, phi, P = _h2_fixture()
    g, b = systems.gpu_system(atoms, shells, nr, na)
    fock = napmo_amd.GridFock(g, b, 10, 'fp64')
    rng = np.random.default_rng(5)
    Q = rng.normal(size=P.shape)
    Q = 1e-2 * (Q + Q.T)
    Fa, _, _ = fock.build(P)
    Fb, _, _ = fock.build(Q)
    Fab, _, _ = fock.build(P + 2.0 * Q)
    assert np.max(np.abs(Fab - (Fa + 2.0 * Fb))) < 1e-12 * np.abs(Fa).max()
    assert np.array_equal(Fa, Fa.T)


def test_grid_fock_chon16_properties():
    # full size: 16 atoms, 50 x 590 per atom (472,000 points), 300 s/p/d functions,
    # lmax 12. Size-independent checks: int rho = tr(P S) = 80, symmetry, trace identity.
    import napmo_amd
    from napmo_amd.wavefunction import grid_ops
    atoms, shells, nr, na, lmax = systems.chon16_system()
    g, b = systems.gpu_system(atoms, shells, nr, na)
    print('chon16 N', g.size, 'n', b['nbasis'])
    assert g.size == 472000 and b['nbasis'] == 300
    fock = napmo_amd.GridFock(g, b, lmax, 'fp64')
    S, _, _ = fock_overlap(fock, b['nbasis'])
    P = systems.density_matrix(np.random.default_rng(2), S, 40)
    F, E, Egrid, rho, v = fock.build(P, want_grid=True)
    print('chon16 E %.12f Egrid %.12f' % (E, Egrid))
    assert abs(g.integrate(rho) - 80.0) < 1e-9
    assert np.array_equal(F, F.T)
    assert abs(E - Egrid) < 1e-9 * abs(E)
    assert E > 0
    # BF16x3 on the same system (one-pass contraction). E_grid comes from the FP64 rho and v
    # and cannot see F; the energy of the F it returns, 1/2 tr(P F), moves by about 2.2e-4 Eh
    # here, so this mode does not meet 1e-6 Eh (the int8 mode does, tests/test_gpu_int8.py)
    mix = napmo_amd.GridFock(g, b, lmax, 'bf16x3')
    F2, E2, Eg2 = mix.build(P)
    print('chon16 bf16x3 max|dF| %.3e dE_grid %.3e dE_trace %.3e' % (np.abs(F2 - F).max(), Eg2 - Egrid, E2 - E))
    assert abs(Eg2 - Egrid) < 1e-9
    assert abs(E2 - E) < 1e-3
    assert np.abs(F2 - F).max() < 2e-5 * np.abs(F).max()


def fock_overlap(fock, n):
    """S = Phi^T diag(w bw) Phi from the same plan: the F of a unit potential is obtained
    by the legacy contraction on the plan's grid."""
    import napmo_amd
    from napmo_amd.wavefunction import grid_ops
    g, b = fock.grid, fock.basis
    phi = b.compute(g.points)
    psi = napmo_amd.PSIB(n)
    S = grid_ops.compute_coupling(psi, g, np.ascontiguousarray(phi.T), np.ones(g.size))
    return S, phi, psi


def test_mfma_lane_layouts():
    # the lane layouts the MFMA kernels assume, checked on one wavefront
    from napmo_amd.system import cext
    for which, K in ((0, 4), (1, 32)):
        out = np.zeros(256)
        cext.check(cext.napmo_selftest_mfma(which, out))
        A = np.array([[m + k / 64.0 for k in range(K)] for m in range(16)])
        B = np.array([[1.0 + (k - n) / 128.0 for n in range(16)] for k in range(K)])
        if which == 1:  # the kernel feeds bf16-rounded operands
            A, B = _bf16_round(A), _bf16_round(B)
        ref = A @ B
        assert np.allclose(out.reshape(16, 16), ref, rtol=1e-6 if which else 1e-14), which


def _bf16_round(x):
    b = np.asarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    r = ((b + 0x7FFF + ((b >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32).astype(np.float64)


def test_grid_fock_orbital_form_matches_density_form():
    import napmo_amd
    atoms, shells, nr, na, o, phi, P = _h2_fixture()
    g, b = systems.gpu_system(atoms, shells, nr, na)
    fock = napmo_amd.GridFock(g, b, 10, 'fp64')
    S = fock.overlap()
    C = systems.density_matrix(np.random.default_rng(1), S, 1)  # P; rebuild C below
    w, V = np.linalg.eigh(C)
    Cocc = V[:, -1:] * np.sqrt(w[-1] / 2.0)                    # P = 2 C C^T
    Fp, Ep, Egp, rhop, vp = fock.build(C, want_grid=True)
    Fo, Eo, Ego, rhoo, vo = fock.build_orbitals(Cocc, 2.0, want_grid=True)
    assert np.max(np.abs(rhoo - rhop)) < 1e-12 * np.abs(rhop).max()
    assert abs(Eo - Ep) < 1e-10 and abs(Ego - Egp) < 1e-10
    assert np.max(np.abs(Fo - Fp)) < 1e-10


def test_orbital_density_streamed_kernel_matches_staged():
    # the orbital-form density with C resident in LDS (Phi streamed, k order permuted inside
    # each 16-function group) against the LDS-staged kernel and the P form, for 1 to 4
    # column blocks of 16 orbitals and a point count that leaves a partial last row tile
    import napmo_amd
    from napmo_amd import autotune
    atoms = _mol(37, natom=5, box=5.0)
    shells = systems.random_shells(np.random.default_rng(37), 5, 14)
    g, b = systems.gpu_system(atoms, shells, 30, 110)
    fock = napmo_amd.GridFock(g, b, 8, 'fp64')
    n = b['nbasis']
    assert n >= 64 and fock.npoint % 48, (n, fock.npoint)
    rng = np.random.default_rng(5)
    for nocc in (1, 17, 40, 64):
        C = rng.standard_normal((n, nocc)) / np.sqrt(n)
        out = {}
        for ds in (0, 1, 2, 3):      # staged kernel, then the three streamed launch shapes
            autotune.apply(fock, dict(autotune.DEFAULTS, density_stream=ds))
            out[ds] = fock.build_orbitals(C, 2.0, want_grid=True)
        Fp, Ep, Egp, rhop, vp = fock.build(2.0 * C @ C.T, want_grid=True)
        r0 = out[0][3]
        for ds in (1, 2, 3):
            rs = out[ds][3]
            print('nocc %d n %d N %d shape %d: max|drho| stream-staged %.2e, stream-P %.2e '
                  '(max rho %.2e)' % (nocc, n, fock.npoint, ds, np.abs(rs - r0).max(),
                                      np.abs(rs - rhop).max(), np.abs(rhop).max()))
            assert np.abs(rs - r0).max() <= 1e-13 * np.abs(r0).max(), (nocc, ds)
            assert np.abs(rs - rhop).max() <= 1e-12 * np.abs(rhop).max(), (nocc, ds)
            assert abs(out[ds][2] - out[0][2]) <= 1e-12 * abs(out[0][2]), (nocc, ds)
            assert np.abs(out[ds][0] - out[0][0]).max() <= 1e-12 * np.abs(out[0][0]).max(), (nocc, ds)
    autotune.apply(fock, autotune.DEFAULTS)


def test_grid_fock_bf16x3_mixed_mode():
    # BF16x3: rho, Poisson and E_grid stay FP64, only F uses the 3-term bf16 split with FP32
    # accumulation. Stated tolerances: max|dF| <= 2e-5 max|F| and |1/2 tr(P dF)| <= 1e-5 |E|
    # (the energy of the returned F; E_grid is FP64 by construction and checked to 1e-10).
    import napmo_amd
    for atoms, shells, nr, na, lmax in (
            (systems.H2_ATOMS, systems.named_shells(systems.H2_ATOMS, 'CC-PVDZ'), 100, 110, 10),
            (_mol(31, natom=5, box=5.0), systems.random_shells(np.random.default_rng(31), 5, 6), 30, 110, 8)):
        g, b = systems.gpu_system(atoms, shells, nr, na)
        f64 = napmo_amd.GridFock(g, b, lmax, 'fp64')
        mix = napmo_amd.GridFock(g, b, lmax, 'bf16x3')
        S = f64.overlap()
        P = systems.density_matrix(np.random.default_rng(3), S, 2)
        F1, E1, Eg1 = f64.build(P)
        F2, E2, Eg2 = mix.build(P)
        dF = np.max(np.abs(F2 - F1))
        print('bf16x3 n=%d max|dF|=%.3e max|F|=%.3e dE_trace=%.3e dE_grid=%.3e' %
              (b['nbasis'], dF, np.abs(F1).max(), E2 - E1, Eg2 - Eg1))
        assert abs(Eg2 - Eg1) < 1e-10
        assert abs(E2 - E1) <= 1e-5 * abs(E1)
        assert dF < 2e-5 * np.abs(F1).max()
        assert np.array_equal(F2, F2.T)


def test_dot_multi_moments_matches_reference_vectors():
    # dot_multi_moments keeps the reference's accumulation order: identical to the last bit
    from napmo_amd.wavefunction import grid_ops
    for mtype in (1, 2, 3, 4):
        out = GOLD['mm%d_init' % mtype].copy()
        grid_ops.dot_multi_moments(GOLD['mm_data'], GOLD['mm_points'], GOLD['mm_center'], 5, mtype,
                                   np.array([20, 7, 30]), out)
        err = np.abs(out - GOLD['mm%d_out' % mtype]).max()
        print('dot_multi_moments mtype', mtype, 'max diff', err)
        assert np.array_equal(out, GOLD['mm%d_out' % mtype]), (mtype, err)
    out = np.zeros((4, 16))
    grid_ops.dot_multi_moments(GOLD['mm_data'][0], GOLD['mm_points'], GOLD['mm_center'], 3, 2,
                               GOLD['mm_empty_seg'], out)
    assert np.array_equal(out, GOLD['mm_empty_out'])
    # the spherical-expansion shape: 40 shells of 110 points, lmax 12, against the oracle
    rng = np.random.default_rng(17)
    pts = rng.normal(size=(4400, 3))
    f = rng.normal(size=4400)
    out = np.zeros((40, 169))
    grid_ops.dot_multi_moments(f, pts, np.zeros(3), 12, 4, np.full(40, 110), out)
    ref = core.dot_multi_moments(f, pts, np.zeros(3), 12, 4, np.full(40, 110), np.zeros((40, 169)))
    assert np.array_equal(out, ref)


def test_density_from_dm_matches_oracle():
    from napmo_amd.wavefunction import grid_ops
    # H2 cc-pVDZ on 2 x 30 x 50 points, dense (epsilon = 0) and screened
    atoms, shells, _, _ = _h2_system()
    g, b = systems.gpu_system(atoms, shells, 30, 50)
    o, ob = systems.oracle_system(atoms, shells, 30, 50)
    phi = ocx.basis_compute(ob, o.points)
    n = phi.shape[1]
    rng = np.random.default_rng(31)
    dm = rng.normal(size=(n, n))
    dmm = np.abs(np.tril(dm)).max(axis=1)
    for eps in (0.0, 1e-3, 100.0):
        out = grid_ops.compute_density_from_dm(b, g, dm, eps, dmm, np.full(g.size, -1.0))
        ref = core.density_from_dm_screened(phi, dm, eps, dmm, np.full(g.size, -1.0))
        print('density_from_dm eps', eps, 'written', int((ref != -1.0).sum()))
        assert np.array_equal(out == -1.0, ref == -1.0)
        assert np.max(np.abs(out - ref)) < 1e-13 * np.abs(ref).max()
    # one centre, diffuse s and p shells: the running threshold stops mid-grid
    atoms = [('H', np.zeros(3))]
    shells = [[{'l': 0, 'prim': [0.05], 'cont': [1.0]}, {'l': 1, 'prim': [0.02], 'cont': [1.0]}]]
    g, b = systems.gpu_system(atoms, shells, 20, 26)
    o, ob = systems.oracle_system(atoms, shells, 20, 26)
    phi = ocx.basis_compute(ob, o.points)
    dm = np.random.default_rng(3).normal(size=(4, 4))
    dmm = np.abs(np.tril(dm)).max(axis=1)
    for eps in (1e-8, 1e-3):
        out = grid_ops.compute_density_from_dm(b, g, dm, eps, dmm, np.full(g.size, -1.0))
        ref = core.density_from_dm_screened(phi, dm, eps, dmm, np.full(g.size, -1.0))
        written = int((ref != -1.0).sum())
        print('density_from_dm one centre eps', eps, 'written', written, 'of', g.size)
        assert 0 < written < g.size
        assert np.array_equal(out == -1.0, ref == -1.0)
        assert np.max(np.abs(out - ref)) < 1e-13 * np.abs(ref).max()


def test_contract_signed_potential_every_kernel():
    # F = Phi^T diag(w bw v) Phi for a potential with both signs (as the multi-species
    # v_eff = V_e - V_p has): the FP64 kernel against the oracle, and every BF16x3 kernel
    # (per-tile, one-pass, and the one-pass form on G = sqrt|s| Phi with the sign applied
    # per point) within 2e-5 of sum_p |w bw v| |phi_mu phi_nu|
    import napmo_amd
    from napmo_amd import autotune
    atoms = _mol(37, natom=5, box=5.0)
    shells = systems.random_shells(np.random.default_rng(37), 5, 6)
    g, b = systems.gpu_system(atoms, shells, 30, 110)
    o, ob = systems.oracle_system(atoms, shells, 30, 110)
    phi = ocx.basis_compute(ob, o.points)
    p = o.points
    v = np.cos(1.3 * p[:, 0]) * np.exp(-0.05 * (p ** 2).sum(1)) - 0.2 * np.sin(p[:, 1])
    assert (v < 0).any() and (v > 0).any()
    ref = core.coupling_matrix(phi, v, o.becke_weights, o.weights)
    scale = core.coupling_matrix(np.abs(phi), np.abs(v), o.becke_weights, o.weights)
    f64 = napmo_amd.GridFock(g, b, 6, 'fp64')
    F = f64.contract(v)
    print('contract fp64 max|dF| %.3e max|F| %.3e' % (np.abs(F - ref).max(), np.abs(ref).max()))
    assert np.abs(F - ref).max() < 1e-13 * np.abs(scale).max()
    mix = napmo_amd.GridFock(g, b, 6, 'bf16x3')
    for use_pass in (0, 1, 2):
        autotune.apply(mix, dict(autotune.DEFAULTS, use_pass=use_pass))
        Fm = mix.contract(v)
        err = np.abs(Fm - ref).max()
        print('contract bf16x3 use_pass=%d max|dF| %.3e scale %.3e' % (use_pass, err, np.abs(scale).max()))
        assert err < 2e-5 * np.abs(scale).max()
        assert np.array_equal(Fm, Fm.T)
