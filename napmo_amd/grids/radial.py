"""Radial integration grid (napmo/grids/radial.py, src/grid/radial.cpp:14-51)."""
import numpy as np

from napmo_amd.grids.radial_transform import ChebyshevRadialTransform
from napmo_amd.system import cext
from napmo_amd.system.elements import atomic_radius


class RadialGrid(object):
    """r_i = r(i), w_i = 4 pi r_i^2 dr/dt (unit t-weights)."""

    def __init__(self, size=2, atomic_symbol='--', rtransform=None):
        self._symbol = atomic_symbol
        radii = atomic_radius(atomic_symbol)
        self._rtransform = rtransform
        if rtransform is None:
            self._rtransform = ChebyshevRadialTransform(radii, size)
        self._this = cext.RadialGrid_new(self._rtransform._this, self._rtransform.radii)

    def __del__(self):
        if getattr(self, '_this', None):
            cext.RadialGrid_del(self._this)
            self._this = None

    def integrate(self, f, segments):
        return cext.RadialGrid_integrate(self._this, segments, np.ascontiguousarray(f, dtype=np.float64))

    def deriv2(self, f):
        """(2/r) f' + f'' by finite differences (src/grid/radial.cpp:53-192)."""
        ptr = cext.RadialGrid_deriv2(self._this, np.ascontiguousarray(f, dtype=np.float64))
        return cext.take_array(ptr, (self.size,))

    @property
    def size(self):
        return cext.RadialGrid_get_size(self._this)

    @property
    def symbol(self):
        return self._symbol

    @property
    def radii(self):
        return cext.RadialGrid_get_radii(self._this)

    @property
    def rtransform(self):
        return self._rtransform

    @property
    def points(self):
        return cext.as_array(cext.RadialGrid_get_points(self._this), (self.size,))

    @property
    def weights(self):
        return cext.as_array(cext.RadialGrid_get_weights(self._this), (self.size,))

    def zeros(self):
        return np.zeros(self.size)
