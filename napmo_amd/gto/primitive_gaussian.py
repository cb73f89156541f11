"""Cartesian primitive Gaussian (napmo/gto/primitive_gaussian.py, src/system/gto.cpp:17-129).
Values and overlaps come from the library (PrimitiveGaussian_compute / _overlap)."""
import numpy as np

from napmo_amd.system import cext


class PrimitiveGaussian(object):

    def __init__(self, exponent=0.5, coefficient=1.0, l=np.zeros(3, dtype=np.int32),
                 origin=np.zeros(3, dtype=np.float64)):
        self._this = cext.PrimitiveGaussian_new(np.ascontiguousarray(l, dtype=np.int32),
                                                np.ascontiguousarray(origin, dtype=np.float64),
                                                float(exponent), float(coefficient))

    @property
    def normalization(self):
        return cext.PrimitiveGaussian_get_norma(self._this)

    @property
    def l(self):
        out = np.zeros(3, dtype=np.int32)
        cext.PrimitiveGaussian_get_l(self._this, out)
        return out

    @property
    def origin(self):
        out = np.zeros(3, dtype=np.float64)
        cext.PrimitiveGaussian_get_origin(self._this, out)
        return out

    @property
    def exponent(self):
        return cext.PrimitiveGaussian_get_zeta(self._this)

    @property
    def coefficient(self):
        return cext.PrimitiveGaussian_get_coeff(self._this)

    def compute(self, coord):
        """Values at coord [n x 3] (napmo/gto/primitive_gaussian.py:56-66)."""
        coord = np.ascontiguousarray(np.atleast_2d(coord), dtype=np.float64)
        out = np.empty(coord.shape[0])
        cext.PrimitiveGaussian_compute(self._this, coord, out, coord.shape[0])
        return out

    def overlap(self, other):
        """<self|other> including coefficients and norms (src/ints/overlap.cpp)."""
        return cext.PrimitiveGaussian_overlap(self._this, other._this)
