"""Cartesian primitive Gaussian (napmo/gto/primitive_gaussian.py, src/system/gto.cpp:17-51)."""
import numpy as np

from napmo_amd.system import cext


class PrimitiveGaussian(object):

    def __init__(self, exponent=0.5, coefficient=1.0, l=np.zeros(3, dtype=np.int32),
                 origin=np.zeros(3, dtype=np.float64)):
        self.exponent = float(exponent)
        self.coefficient = float(coefficient)
        self.l = np.ascontiguousarray(l, dtype=np.int32)
        self.origin = np.ascontiguousarray(origin, dtype=np.float64)
        self._this = cext.PrimitiveGaussian_new(self.l, self.origin, self.exponent, self.coefficient)

    @property
    def normalization(self):
        return cext.PrimitiveGaussian_get_norma(self._this)

    def compute(self, coord):
        coord = np.atleast_2d(coord)
        d = coord - self.origin
        r2 = (d * d).sum(axis=1)
        ang = np.prod(d ** self.l, axis=1)
        return self.coefficient * self.normalization * ang * np.exp(-self.exponent * r2)
