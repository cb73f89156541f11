"""Pins of the oracle's restatement of the screened orbital-form density
(oracle.core.density_orbitals_screened), which the GPU's streamed kernel follows
(tests/test_gpu_density_screen.py):

* with a threshold no block reaches, it is the plain density eta sum_c (Phi C)_pc^2;
* the (16-point tile, 16-function group) products it leaves out at 2^-S move each X_pc by at
  most (groups) 2^-S, so rho_p moves by at most eta sum_c (2 |X_pc| G 2^-S + (G 2^-S)^2);
  checked where a real share of the blocks is left out;
* exponents taken 2^6 too small (a planted mistake) break that bound."""
import numpy as np

from oracle import core
from tests.test_oracle_int8 import _local_system


def _bound(X, G, S, eta):
    return eta * (2 * np.abs(X).sum(1) * G * 2.0 ** -S + X.shape[1] * (G * 2.0 ** -S) ** 2)


def test_screened_density_order_and_bound():
    phi, _, _ = _local_system()                  # 1,920 points, 72 functions, 3 centres
    rng = np.random.default_rng(4)
    C = rng.normal(size=(phi.shape[1], 6)) * 0.5
    eta = 2.0
    X = phi @ C
    full = eta * (X ** 2).sum(1)
    G = -(-phi.shape[1] // 16)
    rho, order, kept = core.density_orbitals_screened(phi, C, eta, 2000)
    assert kept == 1.0 and sorted(order) == list(range(len(order)))
    assert np.abs(rho - full).max() <= 1e-13 * full.max()
    for S in (2, 10, 30):
        rho, order, kept = core.density_orbitals_screened(phi, C, eta, S)
        d = np.abs(rho - full)
        print('S %d: %.1f%% of the tile groups kept, max|drho| %.3e' % (S, 100 * kept, d.max()))
        assert kept < 0.9 and d.max() > 0.0
        assert (d <= _bound(X, G, S, eta) + 1e-13 * full.max()).all()
    real = core._pow2_ceil_exp
    try:
        core._pow2_ceil_exp = lambda m: real(m) - 6
        rho, _, _ = core.density_orbitals_screened(phi, C, eta, 4)
    finally:
        core._pow2_ceil_exp = real
    assert (np.abs(rho - full) > _bound(X, G, 4, eta)).any()
