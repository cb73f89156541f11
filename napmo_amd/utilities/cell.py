"""Non-periodic cell stand-in (napmo/utilities/cell.py); the grid path never uses images."""
from napmo_amd.system import cext


class Cell(object):
    def __init__(self, rvecs=None):
        if rvecs is not None and len(rvecs) != 0:
            raise NotImplementedError('periodic cells are outside the grid hot path')
        self._this = cext.Cell_new(None, 0)

    def __del__(self):
        if getattr(self, '_this', None):
            cext.Cell_del(self._this)
            self._this = None
