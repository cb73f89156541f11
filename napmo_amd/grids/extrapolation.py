"""Spline extrapolations (src/horton/extrapolation.cpp:60-183)."""
from napmo_amd.system import cext


class Extrapolation(object):
    def __del__(self):
        if getattr(self, '_this', None):
            cext.Extrapolation_del(self._this)
            self._this = None


class ZeroExtrapolation(Extrapolation):
    def __init__(self):
        self._this = cext.ZeroExtrapolation_new()


class CuspExtrapolation(Extrapolation):
    def __init__(self):
        self._this = cext.CuspExtrapolation_new()


class PowerExtrapolation(Extrapolation):
    def __init__(self, power):
        self.power = float(power)
        self._this = cext.PowerExtrapolation_new(self.power)


class PotentialExtrapolation(Extrapolation):
    """Left amp r^l, right amp / r^(l+1) (Coulomb potential of angular momentum l)."""

    def __init__(self, l):
        if l < 0:
            raise ValueError('The argument l cannot be negative.')
        self.l = int(l)
        self._this = cext.PotentialExtrapolation_new(self.l)
