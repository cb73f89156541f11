"""Spline extrapolations (src/horton/extrapolation.cpp:60-183)."""
from napmo_amd.system import cext


class Extrapolation(object):
    def __del__(self):
        if getattr(self, '_this', None):
            cext.Extrapolation_del(self._this)
            self._this = None

    def prepare(self, cspline):
        """Re-fit the tails to a CubicSpline's end values (src/horton/extrapolation.cpp:11-13)."""
        cext.Extrapolation_prepare(self._this, cspline._this)

    def eval_left(self, x):
        return cext.Extrapolation_eval_left(self._this, x)

    def eval_right(self, x):
        return cext.Extrapolation_eval_right(self._this, x)

    def deriv_left(self, x):
        return cext.Extrapolation_deriv_left(self._this, x)

    def deriv_right(self, x):
        return cext.Extrapolation_deriv_right(self._this, x)

    def has_tail(self):
        return bool(cext.Extrapolation_has_tail(self._this))

    def to_string(self):
        return self.__class__.__name__


class ZeroExtrapolation(Extrapolation):
    def __init__(self):
        self._this = cext.ZeroExtrapolation_new()


class CuspExtrapolation(Extrapolation):
    def __init__(self):
        self._this = cext.CuspExtrapolation_new()


class PowerExtrapolation(Extrapolation):
    def __init__(self, power):
        self._this = cext.PowerExtrapolation_new(float(power))
        self.power = cext.PowerExtrapolation_get_power(self._this)

    def to_string(self):
        return 'PowerExtrapolation %s' % repr(self.power)


class PotentialExtrapolation(Extrapolation):
    """Left amp r^l, right amp / r^(l+1) (Coulomb potential of angular momentum l)."""

    def __init__(self, l):
        if l < 0:
            raise ValueError('The argument l cannot be negative.')
        self.l = int(l)
        self._this = cext.PotentialExtrapolation_new(self.l)

    def to_string(self):
        return 'PotentialExtrapolation %s' % repr(self.l)
