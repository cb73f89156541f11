"""GPU gradients of the grid Coulomb energy E(P, R) through the torch.autograd glue:
dE/dP (the grid Fock matrix) and dE/dR (basis centres moving on a fixed grid), against
the FP64 oracle, central finite differences and the FP64 Fock matrix."""
import numpy as np
import pytest

from oracle import cext as ocx
from oracle import core
from tests import systems

pytestmark = pytest.mark.gpu

# H2 cc-pVDZ on 2 x 100 x 110 points, lmax 10, tilted so every gradient component is
# non-zero
TILTED_H2 = [('H', (0.1, -0.05, 0.7)), ('H', (-0.08, 0.03, -0.68))]


def test_centre_gradient_matches_oracle():
    import napmo_amd
    atoms = TILTED_H2
    shells = systems.named_shells(atoms, 'CC-PVDZ')
    g, b = systems.gpu_system(atoms, shells, 60, 110)
    o, ob = systems.oracle_system(atoms, shells, 60, 110)
    phi = ocx.basis_compute(ob, o.points)
    S = core.coupling_matrix(phi, np.ones(o.size), o.becke_weights, o.weights)
    P = systems.density_matrix(np.random.default_rng(1), S, 1)
    fock = napmo_amd.GridFock(g, b, 10, 'fp64')
    F, E, Eg, rho, v = fock.build(P, want_grid=True)
    grad = fock.gradient(P)
    ref, cen = core.centre_gradient(ob, o.points, P, o.becke_weights * o.weights * v)
    print('gradient', grad, 'oracle', ref)
    assert np.array_equal(fock.centres(), cen)
    assert np.max(np.abs(grad - ref)) < 1e-11 * np.abs(ref).max()


def test_autograd_dP_and_dR_against_finite_differences():
    import torch
    import napmo_amd
    from napmo_amd.autograd import GridEnergyModel, grid_energy
    atoms = TILTED_H2
    shells = systems.named_shells(atoms, 'CC-PVDZ')
    grid = napmo_amd.BeckeGrid(napmo_amd.make_species(atoms), 100, 110)
    model = GridEnergyModel(grid, [s for s, _ in atoms], shells, 10)
    R0 = np.array([o for _, o in atoms])
    plan = model.plan(R0)
    P = systems.density_matrix(np.random.default_rng(1), plan.overlap(), 1)
    # tensors on the GPU when this torch build sees it (device pointers go straight to the
    # library), else on the CPU (host pointers); the arithmetic is on the GPU either way
    dev = torch.device('cuda' if torch.cuda.is_available() else 'cpu')
    print('autograd tensors on', dev)
    Pt = torch.tensor(P, device=dev, requires_grad=True)
    Rt = torch.tensor(R0, device=dev, requires_grad=True)
    E = grid_energy(Pt, Rt, model)
    gP, gR = torch.autograd.grad(E, (Pt, Rt))
    gP, gR = gP.cpu().numpy(), gR.cpu().numpy()
    # dE/dP is the FP64 grid Fock matrix of the same build
    F, _, Eg = napmo_amd.GridFock(grid, model._basis, 10, 'fp64').build(P)
    assert abs(float(E) - Eg) < 1e-14 * abs(Eg)
    assert np.max(np.abs(gP - F)) < 1e-14 * np.abs(F).max()

    def energy(Pn, Rn):
        with torch.no_grad():
            return float(grid_energy(torch.tensor(Pn, device=dev), torch.tensor(Rn, device=dev), model))

    # central differences in P: symmetric random direction scaled 1e-5 (E is quadratic in
    # P, so the difference is exact up to rounding; what remains is the asymmetry of the
    # discrete Poisson operator, 5e-7 relative in the oracle)
    d = np.random.default_rng(4).normal(size=P.shape)
    d = (d + d.T) * 0.5e-5
    fd = (energy(P + d, R0) - energy(P - d, R0)) / 2.0
    an = float(np.sum(gP * d))
    print('dE/dP . dP: fd %.12e an %.12e rel %.2e' % (fd, an, abs(fd - an) / abs(fd)))
    assert abs(fd - an) < 2e-6 * abs(fd)
    # central differences in R, h = 1e-4 bohr (oracle: |fd - an| <= 1.4e-6 with |g| ~ 0.5)
    h = 1e-4
    gmax = np.abs(gR).max()
    for a in range(2):
        for k in range(3):
            Rp, Rm = R0.copy(), R0.copy()
            Rp[a, k] += h
            Rm[a, k] -= h
            fd = (energy(P, Rp) - energy(P, Rm)) / (2 * h)
            print('dE/dR[%d,%d]: fd %.10e an %.10e diff %.2e' % (a, k, fd, gR[a, k], fd - gR[a, k]))
            assert abs(fd - gR[a, k]) < 1e-5 * gmax


def test_centre_gradient_sixteen_atoms():
    # the 16 x 3 centre gradient of the 16-atom system (300 s/p/d functions, lmax 12) on a
    # reduced grid (20 x 50 per atom), against the oracle; central differences are printed
    # only: with exponents up to 1e4 the discrete Poisson operator is far from self-adjoint,
    # so the finite-difference derivative of E differs from the w bw v form by O(10%)
    import napmo_amd
    atoms, shells, _, _, lmax = systems.chon16_system()
    g, b = systems.gpu_system(atoms, shells, 20, 50)
    o, ob = systems.oracle_system(atoms, shells, 20, 50)
    fock = napmo_amd.GridFock(g, b, lmax, 'fp64')
    P = systems.density_matrix(np.random.default_rng(2), fock.overlap(), 40)
    F, E, Eg, rho, v = fock.build(P, want_grid=True)
    gc = fock.gradient(P)
    vo = core.compute_coulomb(o, core.density_from_dm(ocx.basis_compute(ob, o.points), P), lmax)
    ref, cen = core.centre_gradient(ob, o.points, P, o.becke_weights * o.weights * vo)
    print('16-centre gradient: max|dg| %.3e max|g| %.3e' % (np.abs(gc - ref).max(), np.abs(ref).max()))
    assert gc.shape == (16, 3) and np.array_equal(fock.centres(), cen)
    assert np.max(np.abs(gc - ref)) < 1e-9 * np.abs(ref).max()
