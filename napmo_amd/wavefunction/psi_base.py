"""Minimal PSIB-compatible wavefunction record for the grid contractions
(napmo/wavefunction/psi_base.py:18-45): the same ctypes struct layout, with the
matrices the grid calls write (G, J, XC) allocated as numpy arrays."""
import ctypes

import numpy as np

from napmo_amd.system.cext import WaveFunction


class PSIB(WaveFunction):

    def __init__(self, ndim, eta=2.0, kappa=-1.0, x_factor=None):
        super(PSIB, self).__init__()
        self._ndim = ndim
        self._nbasis = ndim
        self._eta = eta
        self._kappa = kappa
        self._x_factor = (kappa / eta) if x_factor is None else x_factor
        for name in ('S', 'X', 'T', 'V', 'H', 'C', 'D', 'L', 'G', 'J', 'XC', 'F'):
            arr = np.zeros([ndim, ndim])
            setattr(self, name, arr)
            setattr(self, '_' + name, arr.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
        self.O = np.zeros(ndim)
        self._O = self.O.ctypes.data_as(ctypes.POINTER(ctypes.c_double))

    @property
    def ndim(self):
        return self._ndim

    @property
    def x_factor(self):
        return self._x_factor
