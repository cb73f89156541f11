"""Kinetic operator on the grid (napmo/wavefunction/nkinetic.py:15-73).

-1/2 laplacian of a function given on the Becke grid: per atom, the spherical expansion
of psi * bw, deriv2 - l(l+1)/r^2 of every (l, m) component on the Chebyshev radial grid,
splined with PotentialExtrapolation(l) and synthesised at every grid point, all on the
GPU with the projection and synthesis kernels of the Poisson solver.
"""
import numpy as np

from napmo_amd.system import cext


def compute_kinetic(grid, psi, lmax):
    T = np.zeros(grid.size, dtype=np.float64)
    cext.check(cext.napmo_compute_kinetic(grid._this, np.ascontiguousarray(psi, dtype=np.float64), lmax, T))
    return T
