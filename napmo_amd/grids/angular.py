"""Lebedev angular grid (napmo/grids/angular.py, src/grid/angular.cpp:12-37)."""
import numpy as np

from napmo_amd.system import cext


class AngularGrid(object):

    def __init__(self, lorder, spherical=False):
        from napmo_amd.grids.lebedev import lebedev_get_order
        lebedev_get_order(lorder)  # KeyError for an unknown rule, before the library call
        self._this = cext.AngularGrid_new(lorder)
        self._spherical = spherical
        if spherical:  # points become (1, theta, phi) (napmo/grids/angular.py:28-29)
            cext.AngularGrid_spherical(self._this)

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

    def spherical_expansion(self, lmax, f):
        """4 pi sum_j aw_j f[k][j] Y_lm(u_j) per row k of f [nrow x lorder] -> [nrow x L]."""
        f = np.ascontiguousarray(np.atleast_2d(f), dtype=np.float64)
        out = np.empty((f.shape[0], (lmax + 1) ** 2))
        cext.AngularGrid_spherical_expansion(self._this, lmax, f.shape[0], f.reshape(-1), out)
        return out

    def eval_expansion(self, lmax, expansion):
        """sum_lm expansion[k][lm] Y_lm(u_j) -> [nrow x lorder]."""
        e = np.ascontiguousarray(np.atleast_2d(expansion), dtype=np.float64)
        out = np.empty(e.shape[0] * self.lorder)
        cext.AngularGrid_eval_expansion(self._this, lmax, e.shape[0], e, out)
        return out.reshape(e.shape[0], self.lorder)

    @property
    def spherical(self):
        return self._spherical
