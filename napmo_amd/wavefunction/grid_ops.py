"""Grid contractions with the reference call signatures
(src/wavefunction/nwavefunction.cpp via napmo/wavefunction/psi_numeric.py:166, 232,
napmo/wavefunction/psi_hybrid.py:103, napmo/wavefunction/psi_analytic.py:270)."""
import ctypes

import numpy as np

from napmo_amd.system import cext


def compute_2body_matrix(psi, grid, phi, J, K):
    """psi.G = sum_p phi_i phi_j J bw w + x_factor * exchange (phi, K: [ndim x N])."""
    cext.nwavefunction_compute_2body_matrix(ctypes.byref(psi), grid._this, np.ascontiguousarray(phi),
                                            np.ascontiguousarray(J, dtype=np.float64), np.ascontiguousarray(K))
    return psi.G


def compute_coupling(psi, grid, phi, other_J, res=None):
    if res is None:
        res = np.zeros([psi.ndim, psi.ndim])
    cext.nwavefunction_compute_coupling(ctypes.byref(psi), grid._this, np.ascontiguousarray(phi),
                                        np.ascontiguousarray(other_J, dtype=np.float64), res)
    return res


def compute_xc_matrix(psi, grid, phi, vrho):
    cext.nwavefunction_compute_xc_matrix(ctypes.byref(psi), grid._this, np.ascontiguousarray(phi),
                                         np.ascontiguousarray(vrho, dtype=np.float64))
    return psi.XC
