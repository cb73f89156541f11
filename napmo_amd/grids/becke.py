"""Becke multicentre grid (napmo/grids/becke.py:19-164, src/grid/becke.cpp).

The Becke weights are computed on the GPU when the grid is built. Beyond the
reference, atomic grids may differ in size: pass ``sizes=[(nrad, nang), ...]`` (one
per centre); with equal sizes the result is the reference's.

``file=`` reads a Lowdin2 grid (napmo/grids/becke.py:42-50): a Fortran unformatted file
with one int64 record (4 x the point count) and one float64 record of the points and
Becke weights in column-major [npoint x 4] order. Such a grid has points and Becke
weights only (weights 1, no atomic grids): integration and the contractions work on it,
the Poisson and kinetic paths do not.
"""
import ctypes

import numpy as np

from napmo_amd.grids.atomic import AtomicGrid
from napmo_amd.system import cext


def read_lowdin2_grid(path):
    """[npoint x 4] C-ordered (x, y, z, Becke weight) from a Lowdin2 grid file."""
    from scipy.io import FortranFile
    with FortranFile(path, 'r') as f:
        size = int(f.read_ints(dtype=np.int64)[0])
        pw = f.read_reals(dtype=np.float64)
    pw = pw.reshape(int(size / 4), 4, order='F')
    return np.ascontiguousarray(pw)


class BeckeGrid(object):

    def __init__(self, species, n_radial=40, n_angular=110, rtransform=None, file=None, ablmax=None,
                 abldep=None, sizes=None):
        assert isinstance(species, dict)
        centers = species.get('particles')
        ncenter = len(centers)
        self._symbol = species.get('symbol')
        if file is not None:
            pw = read_lowdin2_grid(file)
            self.atgrids = []
            self._this = cext.BeckeGrid_from_points(pw, pw.shape[0], ncenter)
            return
        ablmax = 0 if ablmax is None else ablmax
        abldep = 1.0e-6 if abldep is None else abldep
        self._nrad = n_radial
        self._nang = n_angular
        self.atgrids = []
        for k, center in enumerate(centers):
            nr, na = (n_radial, n_angular) if sizes is None else sizes[k]
            self.atgrids.append(AtomicGrid(nr, na, center.get('origin'), center.get('symbol'),
                                           rtransform=rtransform))
        atgrids_ptr = np.array([g._this for g in self.atgrids], dtype=ctypes.c_void_p)
        self._this = cext.BeckeGrid_new(atgrids_ptr, ncenter, ablmax, abldep)

    def __del__(self):
        if getattr(self, '_this', None):
            cext.BeckeGrid_del(self._this)
            self._this = None

    def evaluate_decomposition(self, atom, cubic_splines, output, cell=None):
        """output += sum_lm U_lm(|r - R_atom|) Y_lm at every grid point (GPU synthesis)."""
        from napmo_amd.utilities.cell import Cell
        if cell is None:
            cell = Cell(None)
        splines = np.array([cs._this for cs in cubic_splines], dtype=ctypes.c_void_p)
        cext.eval_decomposition_grid(splines, np.ascontiguousarray(self.origin[atom, :]), output,
                                     np.ascontiguousarray(self.points), cell._this, len(cubic_splines),
                                     self.size)

    def set_tuning(self, key, value):
        """Grid-level launch knobs (napmo_grid_set_tuning): key 5 = synthesis pairs per lane,
        9 = the smallest last tile of the FP64 MFMA many-density synthesis (17: never)."""
        cext.check(cext.napmo_grid_set_tuning(self._this, int(key), int(value)))

    def get_tuning(self, key):
        v = ctypes.c_long()
        cext.check(cext.napmo_grid_get_tuning(self._this, int(key), ctypes.byref(v)))
        return v.value

    def integrate(self, f):
        return cext.BeckeGrid_integrate(self._this, np.ascontiguousarray(f, dtype=np.float64))

    def show(self):
        print('\nGrid Information:', self._symbol)
        print('-' * (18 + len(str(self._symbol))))
        print('Centers: ', self.ncenter)
        print('Size: ', self.size)

    @property
    def becke_weights(self):
        return cext.as_array(cext.BeckeGrid_get_becke_weights(self._this), (self.size,))

    @property
    def points(self):
        return cext.as_array(cext.BeckeGrid_get_points(self._this), (self.size, 3))

    @property
    def weights(self):
        return cext.as_array(cext.BeckeGrid_get_weights(self._this), (self.size,))

    @property
    def origin(self):
        return cext.as_array(cext.BeckeGrid_get_origin(self._this), (self.ncenter, 3))

    @property
    def size(self):
        return cext.BeckeGrid_get_size(self._this)

    @property
    def radii(self):
        return cext.as_array(cext.BeckeGrid_get_radii(self._this), (self.ncenter,))

    @property
    def ncenter(self):
        return cext.BeckeGrid_get_ncenter(self._this)

    @property
    def nrad(self):
        return self._nrad

    @property
    def nang(self):
        return self._nang

    @property
    def symbol(self):
        return self._symbol
