"""GPU derivatives of the grid Coulomb energy E(P, R) = 1/2 rho^T W V rho (W = w bw, V the
discrete Becke-Poisson operator, rho = diag(Phi(R) P Phi(R)^T), grid fixed, basis functions
moving with their centres) through the library and the torch.autograd glue:

* the transposed operator V^T against V by the dot-product identity <s, V r> = <V^T s, r>;
* the gradient kernel against the oracle's restatement, given the same weight;
* dE/dP and dE/dR against central differences: H2 (cc-pVDZ, 2 x 100 x 110) and the 16-atom
  system (300 s/p/d functions, lmax 12, every one of the 16 x 3 components), dP =
  sym(N(0, 1)) 1e-5 and h = 1e-4 bohr (with its h / 2 Richardson partner)."""
import numpy as np
import pytest

from oracle import cext as ocx
from oracle import core
from tests import systems

pytestmark = pytest.mark.gpu

# H2 cc-pVDZ, tilted so every gradient component is non-zero
TILTED_H2 = [('H', (0.1, -0.05, 0.7)), ('H', (-0.08, 0.03, -0.68))]


@pytest.mark.parametrize('case', ['h2', 'uneven3', 'chon16'])
def test_poisson_adjoint_dot_product(case):
    import napmo_amd
    if case == 'h2':
        atoms, shells, sizes, lmax = TILTED_H2, systems.named_shells(TILTED_H2, 'CC-PVDZ'), [(60, 110)] * 2, 10
    elif case == 'uneven3':
        atoms = [('O', (0.0, 0.0, 0.0)), ('H', (1.1, 0.9, 0.0)), ('N', (-1.2, 0.7, 0.6))]
        shells, sizes, lmax = systems.named_shells(atoms, 'STO-3G'), [(40, 110), (25, 50), (33, 194)], 6
    else:
        atoms, shells, _, _, lmax = systems.chon16_system()
        sizes = [(20, 50)] * 16
    g, b = systems.gpu_system(atoms, shells, None, None, sizes=sizes)
    f = napmo_amd.GridFock(g, b, lmax, 'fp64')
    rng = np.random.default_rng(3)
    r = rng.normal(size=g.size) * np.exp(-rng.uniform(0, 3, g.size))
    s = rng.normal(size=g.size)
    Vr = napmo_amd.compute_coulomb(g, r, lmax)
    Vts = f.poisson_adjoint(s)
    lhs, rhs = float(s @ Vr), float(Vts @ r)
    scale = float(np.abs(s) @ np.abs(Vr))
    print('%s: <s, V r> %.15e  <V^T s, r> %.15e  rel %.2e' % (case, lhs, rhs, abs(lhs - rhs) / scale))
    assert abs(lhs - rhs) < 1e-12 * scale


def test_centre_gradient_kernel_matches_oracle():
    # given the same weight (the symmetrised dE/drho from the library), the gradient kernel
    # equals the oracle's restatement
    import napmo_amd
    atoms = TILTED_H2
    shells = systems.named_shells(atoms, 'CC-PVDZ')
    g, b = systems.gpu_system(atoms, shells, 60, 110)
    o, ob = systems.oracle_system(atoms, shells, 60, 110)
    fock = napmo_amd.GridFock(g, b, 10, 'fp64')
    P = systems.density_matrix(np.random.default_rng(1), fock.overlap(), 1)
    F, E, Eg, rho, v = fock.build(P, want_grid=True)
    W = g.becke_weights * g.weights
    ssym = 0.5 * (W * v + fock.poisson_adjoint(W * rho))
    grad = fock.gradient(P)
    ref, cen = core.centre_gradient(ob, o.points, P, ssym)
    print('gradient', grad, 'oracle', ref)
    assert np.array_equal(fock.centres(), cen)
    assert np.max(np.abs(grad - ref)) < 1e-11 * np.abs(ref).max()
    # dE/dP = Phi^T diag(ssym) Phi
    phi = ocx.basis_compute(ob, o.points)
    dP_ref = core.coupling_matrix(phi, ssym / W, o.becke_weights, o.weights)
    dP = fock.energy_dP()
    assert np.max(np.abs(dP - dP_ref)) < 1e-11 * np.abs(dP_ref).max()


def _fd_check(model, P, R0, h, dev, nprint=6):
    import torch
    from napmo_amd.autograd import grid_energy
    Pt = torch.tensor(P, device=dev, requires_grad=True)
    Rt = torch.tensor(R0, device=dev, requires_grad=True)
    E = grid_energy(Pt, Rt, model)
    gP, gR = torch.autograd.grad(E, (Pt, Rt))
    gP, gR = gP.cpu().numpy(), gR.cpu().numpy()

    def energy(Pn, Rn):
        with torch.no_grad():
            return float(grid_energy(torch.tensor(Pn, device=dev), torch.tensor(Rn, device=dev), model))

    # dE/dP along a symmetric random direction (E is quadratic in P: the central difference
    # is exact up to rounding)
    d = np.random.default_rng(4).normal(size=P.shape)
    d = (d + d.T) * 0.5e-5
    fd = (energy(P + d, R0) - energy(P - d, R0)) / 2.0
    an = float(np.sum(gP * d))
    print('dE/dP . dP: fd %.12e an %.12e rel %.2e' % (fd, an, abs(fd - an) / abs(fd)))
    assert abs(fd - an) < 1e-8 * abs(fd)
    # dE/dR: central differences at h and h / 2 and their Richardson combination (fourth
    # order); with exponents up to 1e4 moving over a coarse fixed grid E(R) has large third
    # derivatives, so the second-order difference at h carries visible truncation error
    gmax = np.abs(gR).max()
    worst2 = worst4 = 0.0
    for a in range(R0.shape[0]):
        for k in range(3):
            e = {}
            for t in (-2, -1, 1, 2):
                Rt_ = R0.copy()
                Rt_[a, k] += t * h / 2
                e[t] = energy(P, Rt_)
            fd2 = (e[2] - e[-2]) / (2 * h)
            fd4 = (8 * (e[1] - e[-1]) - (e[2] - e[-2])) / (6 * h)
            worst2 = max(worst2, abs(fd2 - gR[a, k]))
            worst4 = max(worst4, abs(fd4 - gR[a, k]))
            if a * 3 + k < nprint:
                print('dE/dR[%d,%d]: fd(h) %.10e fd4 %.10e an %.10e' % (a, k, fd2, fd4, gR[a, k]))
    print('dE/dR: max|g| %.3e; max|fd(h) - an| %.3e (%.2e), max|fd4 - an| %.3e (%.2e)' %
          (gmax, worst2, worst2 / gmax, worst4, worst4 / gmax))
    assert worst4 < 1e-5 * gmax
    return gP, gR


def test_autograd_h2_against_finite_differences():
    import torch
    import napmo_amd
    from napmo_amd.autograd import GridEnergyModel
    atoms = TILTED_H2
    shells = systems.named_shells(atoms, 'CC-PVDZ')
    grid = napmo_amd.BeckeGrid(napmo_amd.make_species(atoms), 100, 110)
    model = GridEnergyModel(grid, [s for s, _ in atoms], shells, 10)
    R0 = np.array([o for _, o in atoms])
    plan = model.plan(R0)
    P = systems.density_matrix(np.random.default_rng(1), plan.overlap(), 1)
    # device pointers straight to the library when this torch build sees the GPU, else host
    dev = torch.device('cuda' if torch.cuda.is_available() else 'cpu')
    print('autograd tensors on', dev)
    gP, gR = _fd_check(model, P, R0, 1e-4, dev)
    # the build's own F uses W v alone: it differs from dE/dP by the asymmetry of V
    F, _, _ = napmo_amd.GridFock(grid, model._basis, 10, 'fp64').build(P)
    print('max|F - dE/dP| %.3e (max|F| %.3e)' % (np.abs(F - gP).max(), np.abs(F).max()))


def test_autograd_sixteen_atoms_against_finite_differences():
    # the 16-atom system (300 s/p/d functions, exponents up to 1e4, lmax 12) on a reduced
    # grid (20 x 50 per atom): all 16 x 3 components against central differences
    import torch
    import napmo_amd
    from napmo_amd.autograd import GridEnergyModel
    atoms, shells, _, _, lmax = systems.chon16_system()
    grid = napmo_amd.BeckeGrid(napmo_amd.make_species(atoms), 20, 50)
    model = GridEnergyModel(grid, [s for s, _ in atoms], shells, lmax)
    R0 = np.array([o for _, o in atoms])
    plan = model.plan(R0)
    P = systems.density_matrix(np.random.default_rng(2), plan.overlap(), 40)
    dev = torch.device('cuda' if torch.cuda.is_available() else 'cpu')
    gP, gR = _fd_check(model, P, R0, 1e-4, dev, nprint=3)
    assert gR.shape == (16, 3)


@pytest.mark.parametrize('case', ['h2', 'chon16'])
def test_recollocating_backward_matches_stored_phi(case):
    # NAPMO_TUNE_BACKWARD_POINTS > 0: dE/dP and dE/dR from Phi recollocated chunk by chunk
    # (neither the stored Phi nor a full Phi P), bit for bit the stored-Phi results
    import ctypes
    import napmo_amd
    from napmo_amd.system import cext
    if case == 'h2':
        atoms, shells, nr, na, lmax = TILTED_H2, systems.named_shells(TILTED_H2, 'CC-PVDZ'), 60, 110, 10
        P_occ = 1
    else:
        atoms, shells, _, _, lmax = systems.chon16_system()
        nr, na, P_occ = 20, 50, 40
    g, b = systems.gpu_system(atoms, shells, nr, na)
    f = napmo_amd.GridFock(g, b, lmax, 'fp64')
    P = systems.density_matrix(np.random.default_rng(6), f.overlap(), P_occ)
    f.build(P)
    dP0, gR0 = f.energy_dP(), f.gradient(P)
    key = 10  # NAPMO_TUNE_BACKWARD_POINTS
    for pts in (1, 50000):
        cext.check(cext.napmo_fock_plan_set_tuning(f._this, key, pts))
        v = ctypes.c_long()
        cext.check(cext.napmo_fock_plan_get_tuning(f._this, key, ctypes.byref(v)))
        dP1, gR1 = f.energy_dP(), f.gradient(P)
        print('%s: %d points per chunk (%d asked) on %d points: dE/dP and dE/dR equal: %s %s'
              % (case, v.value, pts, g.size, np.array_equal(dP1, dP0), np.array_equal(gR1, gR0)))
        assert np.array_equal(dP1, dP0) and np.array_equal(gR1, gR0)
    cext.check(cext.napmo_fock_plan_set_tuning(f._this, key, 0))
