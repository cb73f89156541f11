"""End-to-end grid Hartree-Fock anchor on two centres: the reference's numerical H2 test
(napmo/wavefunction/test/test_psi_numeric_becke_H2.py:14-95): H2 with the nuclei at
z = +-0.356 Angstrom (1.889726133921252 bohr per Angstrom, napmo/data/constants.py:34), two
100 x 110 Becke atomic grids, lmax 8, started from the STO-3G bonding orbital, optimised by
Becke-Dickson corrections (napmo/wavefunction/psi_optimization.py:124-221, PSIO_BECKE) until
the energy changes by less than 1e-8 Eh. Asserted as the reference asserts (rtol 1e-5 against
-1.117506 for the STO-3G start, rtol 1e-2 against -1.133630 for the grid result), and against
the reference's own grid result on this grid, -1.13306385225865 Eh (its test's docstring), to
1e-5 Eh. Unlike He (one centre, where the same driver meets the reference's grid energy to
7e-9 Eh), the two-centre energy keeps sinking by ~1e-8 Eh per step after the first ten
(-1.133068177 at the 1e-8 stop, -1.133069055 after 60 steps), so the value depends on where
an optimisation stops: the reference stops at |dE| < 1e-7 on its own path (DIIS in its
subspace step), 4.3e-6 Eh above ours; the test asserts ours is the lower.

Every grid operator runs on the GPU through the library: kinetic (napmo_compute_kinetic),
nuclear (napmo_compute_nuclear), Coulomb and exchange potentials (napmo_compute_coulomb) and
the orbital correction (napmo_compute_dpsi), on a two-centre Becke grid. The STO-3G start
needs no integrals library: with one s function per centre the bonding orbital is fixed by
symmetry, and its overlap, kinetic, nuclear and (ss|ss) integrals have closed forms (Gaussian
product theorem, Boys F0), which the grid operators are checked against first."""
import json
import math
import os

import numpy as np
import pytest
import scipy.linalg

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ANGSTROM_TO_BOHR = 1.889726133921252
E_GTO = -1.117506               # the reference test's STO-3G assertion (rtol 1e-5)
E_LIMIT = -1.133630             # its grid assertion target (rtol 1e-2)
E_REF_GRID = -1.13306385225865  # the reference's own 100 x 110 grid result


def _boys0(t):
    return 1.0 if t < 1e-15 else 0.5 * math.sqrt(math.pi / t) * math.erf(math.sqrt(t))


def _sto3g_h():
    with open(os.path.join(ROOT, 'napmo_amd', 'data', 'basis_sets.json')) as f:
        sh = json.load(f)['STO-3G']['H'][0]
    return np.array(sh['prim']), np.array(sh['cont'])


def _closed_form_start(zeta, coef, centres):
    """The bonding orbital of two normalised contracted s functions: its one-electron
    energy h = <T + V_nuc>, Coulomb self energy (gg|gg) and the total RHF energy."""
    prims = [(a, c * (2.0 * a / math.pi) ** 0.75, np.array(R)) for R in centres for a, c in zip(zeta, coef)]
    nb = len(zeta)
    fn = [prims[:nb], prims[nb:]]

    def pair(pa, pb):
        a, ca, A = pa
        b, cb, B = pb
        p = a + b
        P = (a * A + b * B) / p
        return p, P, ca * cb * math.exp(-a * b / p * float((A - B) @ (A - B)))

    def one(f, g):
        S = T = V = 0.0
        for pa in f:
            for pb in g:
                p, P, K = pair(pa, pb)
                mu = pa[0] * pb[0] / p
                R2 = float((pa[2] - pb[2]) @ (pa[2] - pb[2]))
                S += K * (math.pi / p) ** 1.5
                T += K * mu * (3.0 - 2.0 * mu * R2) * (math.pi / p) ** 1.5
                for C in centres:
                    V += -K * 2.0 * math.pi / p * _boys0(p * float((P - np.array(C)) @ (P - np.array(C))))
        return S, T, V

    def eri(f, g, h, k):
        s = 0.0
        for pa in f:
            for pb in g:
                p, P, K1 = pair(pa, pb)
                for pc in h:
                    for pd in k:
                        q, Q, K2 = pair(pc, pd)
                        al = p * q / (p + q)
                        s += K1 * K2 * 2.0 * math.pi ** 2.5 / (p * q * math.sqrt(p + q)) * \
                            _boys0(al * float((P - Q) @ (P - Q)))
        return s

    S = np.zeros((2, 2))
    H = np.zeros((2, 2))
    for i in range(2):
        for j in range(2):
            s, t, v = one(fn[i], fn[j])
            S[i, j], H[i, j] = s, t + v
    nrm = [1.0 / math.sqrt(S[i, i]) for i in range(2)]
    S = S * np.outer(nrm, nrm)
    H = H * np.outer(nrm, nrm)
    cg = 1.0 / math.sqrt(2.0 * (1.0 + S[0, 1]))           # sigma_g = cg (phi_A + phi_B)
    c = np.array([cg, cg])
    h = float(c @ H @ c)
    J = 0.0
    for i in range(2):
        for j in range(2):
            for k in range(2):
                for l in range(2):
                    J += c[i] * c[j] * c[k] * c[l] * nrm[i] * nrm[j] * nrm[k] * nrm[l] * eri(fn[i], fn[j], fn[k], fn[l])
    R = float(np.linalg.norm(np.array(centres[0]) - np.array(centres[1])))
    return cg, nrm, h, J, 2.0 * h + J + 1.0 / R


def test_h2_100x110_grid_hartree_fock():
    import napmo_amd
    z = 0.356 * ANGSTROM_TO_BOHR
    atoms = [('H', (0.0, 0.0, z)), ('H', (0.0, 0.0, -z))]
    grid = napmo_amd.BeckeGrid(napmo_amd.make_species(atoms), 100, 110)
    lmax = 8
    assert grid.size == 22000
    zeta, coef = _sto3g_h()
    cg, nrm, h0, J0, E0 = _closed_form_start(zeta, coef, [o for _, o in atoms])
    print('STO-3G start (closed forms): h %.10f (gg|gg) %.10f E %.10f (reference %.6f)' % (h0, J0, E0, E_GTO))
    assert np.allclose(E0, E_GTO, rtol=1e-5)
    E_nn = 1.0 / (2.0 * z)
    basis = napmo_amd.make_basis(atoms, basis_name='STO-3G')
    phi = basis.compute(grid.points)                        # [N x 2]
    psi = np.ascontiguousarray(cg * (phi[:, 0] + phi[:, 1]))
    I = grid.integrate
    charges = [{'atomic_number': 1, 'origin': o} for _, o in atoms]
    Vn = -napmo_amd.compute_nuclear(grid, charges)
    J = lambda d: napmo_amd.compute_coulomb(grid, d, lmax)      # noqa: E731
    T = lambda f: napmo_amd.compute_kinetic(grid, f, lmax)      # noqa: E731

    # the start against its closed forms
    n2 = I(psi * psi)
    hg = I(psi * T(psi)) + I(psi * Vn * psi)
    Jg = I(psi * J(psi * psi) * psi)
    print('grid start: <psi|psi> %.10f  h %.10f (%.10f)  (gg|gg) %.10f (%.10f)' % (n2, hg, h0, Jg, J0))
    assert abs(n2 - 1.0) < 1e-5
    assert np.allclose([hg, Jg], [h0, J0], rtol=1e-4)
    psi = psi / math.sqrt(n2)

    # Becke-Dickson optimisation of the doubly occupied orbital, as for He
    def fock(b, J1):
        return T(b) + Vn * b + 2.0 * J1 * b - J(psi * b) * psi

    J1 = J(psi * psi)
    eps = I(psi * fock(psi, J1))
    E_prev, E = 0.0, None
    for it in range(1, 101):
        fpsi = fock(psi, J1)
        R = fpsi - eps * psi
        de = I(R * psi)
        V_tot = fpsi - T(psi)
        dpsi = napmo_amd.compute_dpsi(grid, lmax, psi, de, eps, R, V_tot, 1.0)
        basis2 = [psi, dpsi]
        fb = [fock(b, J1) for b in basis2]
        S = np.array([[I(a * b) for b in basis2] for a in basis2])
        F = np.array([[I(a * g) for g in fb] for a in basis2])
        F = 0.5 * (F + F.T)
        c = scipy.linalg.eigh(F, S)[1][:, 0]
        psi = c[0] * psi + c[1] * dpsi
        psi = psi / math.sqrt(I(psi * psi))
        J1 = J(psi * psi)
        h = I(psi * (T(psi) + Vn * psi))
        E = 2.0 * h + I(psi * J1 * psi) + E_nn
        eps = I(psi * fock(psi, J1))
        print('%3d  E %.12f  dE %.3e  eps %.10f' % (it, E, E - E_prev, eps))
        if abs(E - E_prev) < 1e-8:
            break
        E_prev = E
    print('H2 2 x 100 x 110: E %.12f; reference grid result %.12f (diff %.2e); limit %.6f' %
          (E, E_REF_GRID, E - E_REF_GRID, E_LIMIT))
    assert abs(E - E_prev) < 1e-8, 'not converged'
    assert np.allclose(E, E_LIMIT, rtol=1e-2)
    # the reference's own grid energy on the same discretisation, reached and passed (the
    # energy is variational in the orbital on a fixed grid)
    assert abs(E - E_REF_GRID) < 1e-5 and E < E_REF_GRID
