"""Atomic grid = radial x Lebedev (napmo/grids/atomic.py, src/grid/atomic_grid.cpp)."""
import numpy as np

from napmo_amd.grids.angular import AngularGrid
from napmo_amd.grids.radial import RadialGrid
from napmo_amd.system import cext


class AtomicGrid(object):

    def __init__(self, nrad, nang, origin, atomic_symbol, rtransform=None):
        self.radial_grid = RadialGrid(nrad, atomic_symbol, rtransform=rtransform)
        self.angular_grid = AngularGrid(nang)
        self._symbol = atomic_symbol
        self._this = cext.AtomicGrid_new(self.angular_grid._this, self.radial_grid._this,
                                         np.ascontiguousarray(origin, dtype=np.float64))

    def __del__(self):
        if getattr(self, '_this', None):
            cext.AtomicGrid_del(self._this)
            self._this = None

    def spherical_expansion(self, lmax, f):
        """f_lm(r_i) per radial shell, shape (nrad, (lmax+1)^2); computed on the GPU."""
        if lmax < 0:
            raise ValueError('lmax can not be negative.')
        lsize = (lmax + 1) * (lmax + 1)
        ptr = cext.AtomicGrid_spherical_expansion(self._this, np.ascontiguousarray(f, dtype=np.float64), lmax)
        return cext.take_array(ptr, (self.radial_grid.size, lsize))

    def evaluate_expansion(self, lmax, expansion):
        """f(r_i, u_j) = sum_lm expansion[i][lm] Y_lm(u_j) for one atom, with the real
        spherical harmonics of AngularGrid_spherical_expansion (napmo/grids/atomic.py:57-75)."""
        output = np.empty(self.size, dtype=np.float64)
        cext.AngularGrid_eval_expansion(self.angular_grid._this, lmax, self.radial_grid.size,
                                        np.ascontiguousarray(expansion, dtype=np.float64), output)
        return output

    def spherical_average(self, *args):
        f = np.concatenate(args)
        s = self.integrate(segmented=True)
        output = self.integrate(f, segmented=True)
        output /= s
        return output

    def integrate(self, *args, **kwargs):
        segmented = kwargs.pop('segmented', False)
        args += (self.weights,)
        nfunc = len(args)
        if segmented:
            nseg = self.radial_grid.size
            sseg = self.angular_grid.lorder
        else:
            nseg = 1
            sseg = self.size
        f = np.ascontiguousarray(np.concatenate(args), dtype=np.float64)
        ptr = cext.AtomicGrid_integrate(self._this, nfunc, nseg, sseg, f)
        return cext.take_array(ptr, (nseg,))

    @property
    def points(self):
        return cext.as_array(cext.AtomicGrid_get_points(self._this), (self.size, 3))

    @property
    def weights(self):
        return cext.as_array(cext.AtomicGrid_get_weights(self._this), (self.size,))

    @property
    def origin(self):
        return cext.as_array(cext.AtomicGrid_get_origin(self._this), (3,))

    @property
    def size(self):
        return cext.AtomicGrid_get_size(self._this)

    @property
    def radii(self):
        return cext.AtomicGrid_get_radii(self._this)

    @property
    def symbol(self):
        return self._symbol
