"""Nuclear attraction potential on the grid (napmo/wavefunction/nnuclear.py:15-37)."""
import numpy as np

from napmo_amd.system import cext


def compute_nuclear(grid, point_charges):
    """sum_A Z_A / |r_i - R_A| at every grid point; point_charges: dicts with
    'atomic_number' and 'origin' (the reference's keys)."""
    q = np.array([p.get('atomic_number', None) for p in point_charges], dtype=np.float64)
    o = np.ascontiguousarray(np.array([p.get('origin') for p in point_charges], dtype=np.float64).reshape(-1, 3))
    V = np.zeros(grid.size, dtype=np.float64)
    cext.check(cext.napmo_compute_nuclear(grid._this, len(q), q, o, V))
    return V
