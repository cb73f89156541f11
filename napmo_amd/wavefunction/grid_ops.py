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


def compute_density_from_dm(basis, grid, dm, epsilon=0.0, dmmaxrow=None, out=None):
    """rho on the grid's points from the lower triangle of dm (nwavefunction.cpp:109-159,
    bound at napmo/__init__.py:182-184). epsilon > 0 applies the reference's screening,
    which may stop early and leave the tail of ``out`` as it was."""
    n = basis.get('nbasis')
    if out is None:
        out = np.zeros(grid.size)
    if dmmaxrow is None:
        dmmaxrow = np.abs(np.asarray(dm)).max(axis=1)
    cext.nwavefunction_compute_density_from_dm(basis._this, grid._this,
                                               np.ascontiguousarray(dm, dtype=np.float64).reshape(n, n),
                                               out, float(epsilon),
                                               np.ascontiguousarray(dmmaxrow, dtype=np.float64))
    return out


def dot_multi_moments(data, points, center, lmax, mtype, segments, output):
    """src/horton/dot_multi.cpp:148-186 on the GPU: output[seg, m] += sum over the segment's
    points of prod_v data[v, p] * poly_m(points[p] - center). data [nvector x npoint]."""
    data = np.ascontiguousarray(np.atleast_2d(data), dtype=np.float64)
    nvector, npoint = data.shape
    cext.dot_multi_moments(npoint, nvector, data.reshape(-1), np.ascontiguousarray(points, dtype=np.float64),
                           np.ascontiguousarray(center, dtype=np.float64), lmax, mtype,
                           np.ascontiguousarray(segments, dtype=np.int64), output, output.shape[-1])
    return output
