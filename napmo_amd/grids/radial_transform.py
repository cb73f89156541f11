"""Radial transforms r(t) on the uniform t-axis 0..n-1 (napmo/grids/radial_transform.py).

Thin objects over library handles; the formulas are those of
src/horton/rtransform.cpp:257-283 (power) and :308-341 (Chebyshev).
"""
import numbers

import numpy as np

from napmo_amd.system import cext


class RadialTransform(object):
    """Base class; see napmo/grids/radial_transform.py:14-105."""

    def __init__(self):
        self._size = cext.RTransform_get_npoint(self._this)
        self._points = np.arange(self.size, dtype=np.float64)

    def __del__(self):
        if getattr(self, '_this', None):
            cext.RTransform_del(self._this)
            self._this = None

    def _eval(self, t, scalar, array):
        if isinstance(t, numbers.Number):
            return scalar(self._this, t)
        if isinstance(t, np.ndarray):
            return self._all(array, t)
        raise NotImplementedError

    def _all(self, fn, points=None):
        if points is None:
            points = self._points
        points = np.ascontiguousarray(points, dtype=np.float64)
        data = np.zeros(points.shape[0], dtype=np.float64)
        fn(self._this, points, data, points.shape[0])
        return data

    def radius(self, t):
        return self._eval(t, cext.RTransform_radius, cext.RTransform_radius_array)

    def deriv(self, t):
        return self._eval(t, cext.RTransform_deriv, cext.RTransform_deriv_array)

    def deriv2(self, t):
        return self._eval(t, cext.RTransform_deriv2, cext.RTransform_deriv2_array)

    def deriv3(self, t):
        return self._eval(t, cext.RTransform_deriv3, cext.RTransform_deriv3_array)

    def inv(self, r):
        return self._eval(r, cext.RTransform_inv, cext.RTransform_inv_array)

    def radius_all(self, points=None):
        return self._all(cext.RTransform_radius_array, points)

    def deriv_all(self, points=None):
        return self._all(cext.RTransform_deriv_array, points)

    def deriv2_all(self, points=None):
        return self._all(cext.RTransform_deriv2_array, points)

    def deriv3_all(self, points=None):
        return self._all(cext.RTransform_deriv3_array, points)

    def inv_all(self, points):
        return self._all(cext.RTransform_inv_array, points)

    def get_radius(self):
        return self.radius_all()

    @property
    def size(self):
        return self._size


class IndentityRadialTransform(RadialTransform):
    """r = t (testing only; the reference's spelling is kept)."""

    def __init__(self, size):
        self._this = cext.IdentityRTransform_new(size)
        super(IndentityRadialTransform, self).__init__()

    def to_string(self):
        return ' '.join(['IdentityRTransform', repr(self.size)])


class PowerRadialTransform(RadialTransform):
    """r_i = rmin (i+1)^power, power = ln(rmax/rmin) / ln(n)."""

    def __init__(self, rmin, rmax, size):
        if size < 2:
            raise ValueError('A radial grid consists of at least two points.')
        if rmin >= rmax:
            raise ValueError('rmin must be below rmax.')
        self._this = cext.PowerRTransform_new(rmin, rmax, size)
        self._rmin = cext.PowerRTransform_get_rmin(self._this)
        self._rmax = cext.PowerRTransform_get_rmax(self._this)
        self._power = cext.PowerRTransform_get_power(self._this)
        super(PowerRadialTransform, self).__init__()

    def to_string(self):
        return ' '.join(['PowerRTransform', repr(self.rmin), repr(self.rmax), repr(self.size)])

    @property
    def rmin(self):
        return self._rmin

    @property
    def rmax(self):
        return self._rmax

    @property
    def power(self):
        return self._power

    @property
    def radii(self):
        return self._rmax


class ChebyshevRadialTransform(RadialTransform):
    """Becke's r = radii (1 + x) / (1 - x), x = cos(pi (n - t) / (n + 1))."""

    def __init__(self, radii, size):
        if size < 2:
            raise ValueError('A radial grid consists of at least two points.')
        self._this = cext.ChebyshevRTransform_new(radii, size)
        self._radii = cext.ChebyshevRTransform_get_radii(self._this)
        super(ChebyshevRadialTransform, self).__init__()

    def to_string(self):
        return ' '.join(['ChebyshevRTransform', repr(self.radii), repr(self.size)])

    @property
    def radii(self):
        return self._radii

    @property
    def z(self):
        return cext.ChebyshevRTransform_get_z(self._this)
