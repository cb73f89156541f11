"""Cubic spline on a transformed radial axis (napmo/grids/cubic_spline.py:16-140)."""
import numpy as np

from napmo_amd.grids.extrapolation import CuspExtrapolation
from napmo_amd.grids.radial_transform import IndentityRadialTransform
from napmo_amd.system import cext


class CubicSpline(object):

    def __init__(self, y, dx=None, rtransform=None, extrapolation=None, v=None):
        self._y = np.ascontiguousarray(y, dtype=np.float64)
        n = self._y.shape[0]
        if rtransform is None:
            self._rtransform = IndentityRadialTransform(n)
        else:
            self._rtransform = rtransform
            assert rtransform.size == n
        if v is None:
            v = self._rtransform.deriv_all()
        if dx is None:
            self._dt = np.zeros(n, np.float64)
            cext.solve_cubic_spline_system(self._y, self._dt, n)
            self._dx = self._dt / v
        else:
            assert dx.shape[0] == n
            self._dx = dx
            self._dt = np.ascontiguousarray(dx * v, dtype=np.float64)
        self._extrapolation = CuspExtrapolation() if extrapolation is None else extrapolation
        self._this = cext.CubicSpline_new(self._y, self._dt, self._extrapolation._this,
                                          self._rtransform._this, n)

    @classmethod
    def from_dt(cls, y, dt, rtransform, extrapolation):
        """Spline from t-derivatives directly (used for the Poisson solutions)."""
        obj = cls.__new__(cls)
        obj._y = np.ascontiguousarray(y, dtype=np.float64)
        obj._dt = np.ascontiguousarray(dt, dtype=np.float64)
        obj._rtransform = rtransform
        obj._dx = obj._dt / rtransform.deriv_all()
        obj._extrapolation = extrapolation
        obj._this = cext.CubicSpline_new(obj._y, obj._dt, extrapolation._this, rtransform._this,
                                         obj._y.shape[0])
        return obj

    def __del__(self):
        if getattr(self, '_this', None):
            cext.CubicSpline_del(self._this)
            self._this = None

    def __call__(self, new_x, new_y=None):
        new_x = np.ascontiguousarray(new_x, dtype=np.float64)
        new_n = new_x.shape[0]
        if new_y is None:
            new_y = np.zeros(new_n, np.float64)
        cext.CubicSpline_eval(self._this, new_x, new_y, new_n)
        return new_y

    def deriv(self, new_x, new_dx=None):
        """dy/dx at new_x (napmo/grids/cubic_spline.py:96-117)."""
        new_x = np.ascontiguousarray(new_x, dtype=np.float64)
        new_n = new_x.shape[0]
        if new_dx is None:
            new_dx = np.zeros(new_n, np.float64)
        else:
            assert new_dx.shape[0] == new_n
        cext.CubicSpline_eval_deriv(self._this, new_x, new_dx, new_n)
        return new_dx

    @property
    def y(self):
        return self._y

    @property
    def dx(self):
        return self._dx

    @property
    def dt(self):
        return self._dt

    @property
    def first_x(self):
        return cext.CubicSpline_get_first_x(self._this)

    @property
    def last_x(self):
        return cext.CubicSpline_get_last_x(self._this)

    @property
    def extrapolation(self):
        return self._extrapolation

    @property
    def rtransform(self):
        return self._rtransform
