"""Grid exchange potential (napmo/wavefunction/psi_numeric.py:186-215) on the GPU.

aux_sv = V[psi_s psi_v] for every s >= v by the Becke-Poisson chain (one device
expansion and synthesis per pair, with the plan's precomputed Green matrices), then
Kpot[j] = sum_l aux_lj (D^T psi)_l, accumulated on the device.
"""
import numpy as np

from napmo_amd.system import cext


def compute_exchange_potential(grid, psi, D, lmax):
    """psi [ndim x N] on the grid, D [ndim x ndim]; returns Kpot [ndim x N]."""
    psi = np.ascontiguousarray(psi, dtype=np.float64)
    ndim = psi.shape[0]
    K = np.zeros_like(psi)
    cext.check(cext.napmo_compute_exchange_potential(grid._this, ndim, psi,
                                                     np.ascontiguousarray(D, dtype=np.float64), lmax, K))
    return K
