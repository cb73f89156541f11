"""Lebedev angular grid (napmo/grids/angular.py, src/grid/angular.cpp:12-37)."""
import numpy as np

from napmo_amd.system import cext


class AngularGrid(object):

    def __init__(self, lorder, spherical=False):
        if spherical:
            raise NotImplementedError('spherical-coordinate angular grids are not on the hot path')
        from napmo_amd.grids.lebedev import lebedev_get_order
        lebedev_get_order(lorder)  # KeyError for an unknown rule, before the library call
        self._this = cext.AngularGrid_new(lorder)
        self._spherical = spherical

    def __del__(self):
        if getattr(self, '_this', None):
            cext.AngularGrid_del(self._this)
            self._this = None

    def integrate(self, *args):
        f = np.ascontiguousarray(np.concatenate(args), dtype=np.float64)
        return cext.AngularGrid_integrate(self._this, len(args), f)

    @property
    def lorder(self):
        return cext.AngularGrid_get_lorder(self._this)

    @property
    def points(self):
        return cext.as_array(cext.AngularGrid_get_points(self._this), (self.lorder, 3))

    @property
    def weights(self):
        return cext.as_array(cext.AngularGrid_get_weights(self._this), (self.lorder,))

    @property
    def spherical(self):
        return self._spherical
