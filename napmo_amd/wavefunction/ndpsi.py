"""Becke-Dickson orbital correction on the GPU (napmo/wavefunction/ndpsi.py:15-89) and the
per-orbital steps of PSIO_BECKE.optimize (napmo/wavefunction/psi_optimization.py:175-203).

For every orbital, atom and (l, m) the correction solves the radial equation
u'' + (2/r) u' + a u = -2/m_inv mu_lm with a = -2/m_inv (l(l+1)/r^2 + V_sph - oi - doi) and
u = 0 at both ends of the atom's radial grid; mu_lm is the spherical expansion of
(phi doi - ri) bw and V_sph the spherical average of V_tot bw. The matrix depends on the
orbital, so each (orbital, atom, l) system is assembled, band-factored and solved on the
device; the solutions are synthesised at every grid point with the Poisson kernels.
"""
import numpy as np

from napmo_amd.system import cext


def _rows(a, norb, n):
    a = np.ascontiguousarray(a, dtype=np.float64).reshape(norb, n)
    return a


def compute_dpsi(grid, lmax, phi, doi, oi, ri, V_tot, mass_inv):
    """The reference's signature: one orbital phi [N], floats doi, oi, arrays ri, V_tot [N].
    Several orbitals at once: phi, ri, V_tot [norb x N] and doi, oi [norb]."""
    phi = np.asarray(phi, dtype=np.float64)
    single = phi.ndim == 1
    norb = 1 if single else phi.shape[0]
    n = grid.size
    phi = _rows(phi, norb, n)
    ri = _rows(ri, norb, n)
    V = _rows(V_tot, norb, n)
    doi = np.ascontiguousarray(np.broadcast_to(np.asarray(doi, dtype=np.float64), (norb,)))
    oi = np.ascontiguousarray(np.broadcast_to(np.asarray(oi, dtype=np.float64), (norb,)))
    out = np.zeros((norb, n))
    cext.check(cext.napmo_compute_dpsi(grid._this, int(lmax), norb, phi.ctypes.data, doi.ctypes.data,
                                       oi.ctypes.data, ri.ctypes.data, V.ctypes.data, float(mass_inv),
                                       out.ctypes.data))
    return out[0] if single else out


def becke_dickson_step(grid, lmax, psi, T, V_tot, e_prev, mass_inv):
    """Residual, energy correction and orbital correction of every orbital:
    psi, T (kinetic operator applied), V_tot (potential applied) [norb x N], e_prev [norb].
    Returns (dpsi [norb x N], delta_e [norb], R [norb x N])."""
    psi = np.asarray(psi, dtype=np.float64)
    norb = psi.shape[0]
    n = grid.size
    psi, T, V = _rows(psi, norb, n), _rows(T, norb, n), _rows(V_tot, norb, n)
    e = np.ascontiguousarray(e_prev, dtype=np.float64).reshape(norb)
    dpsi, de, R = np.zeros((norb, n)), np.zeros(norb), np.zeros((norb, n))
    cext.check(cext.napmo_becke_dickson_step(grid._this, int(lmax), norb, psi.ctypes.data,
                                             T.ctypes.data, V.ctypes.data, e.ctypes.data,
                                             float(mass_inv), dpsi.ctypes.data, de.ctypes.data,
                                             R.ctypes.data))
    return dpsi, de, R


def becke_dickson_step_device(grid, lmax, norb, d_psi, d_T, d_V, d_e, mass_inv, d_dpsi, d_de, d_R):
    """Same on device pointers (ints) resident on the grid's GPU; returns when done."""
    cext.check(cext.napmo_becke_dickson_step_device(grid._this, int(lmax), int(norb), d_psi, d_T,
                                                    d_V, d_e, float(mass_inv), d_dpsi, d_de, d_R))
