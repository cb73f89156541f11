"""Several species' grids and their common points (napmo/grids/multi_grid.py:13-146,
src/grid/multi_grid.cpp). The common-point index is built by hashing the smaller grid's
coordinates (exact equality, as the reference's pairwise comparison)."""
import numpy as np

from napmo_amd.system import cext


class MultiGrid(object):

    def __init__(self, nspecies):
        assert isinstance(nspecies, int)
        self._grids = {}
        self._this = cext.MultiGrid_new(nspecies)

    def __del__(self):
        if getattr(self, '_this', None):
            cext.MultiGrid_del(self._this)
            self._this = None

    def add_grid(self, grid):
        from napmo_amd.grids.becke import BeckeGrid
        assert isinstance(grid, BeckeGrid)
        gid = cext.MultiGrid_add_grid(self._this, grid._this)
        self._grids[grid._symbol] = {'id': gid, 'obj': grid}

    def get_grid(self, symbol='', gid=-1):
        assert isinstance(symbol, str)
        assert isinstance(gid, int)
        aux = self._grids.get(symbol, None)
        if aux:
            return aux.get('obj', None)
        for grid in self._grids.values():
            if grid.get('id', -2) == gid:
                return grid.get('obj', None)
        return aux

    def get_grid_id(self, symbol):
        assert isinstance(symbol, str)
        aux = self._grids.get(symbol, None)
        if aux:
            return aux.get('id', None)
        return aux

    def get_common_points(self, symbol_a, symbol_b):
        """[n x 2] indices of the points with equal coordinates: column 0 indexes the larger
        grid (symbol_a on a tie), column 1 the other; the same grid gives arange(size)."""
        assert isinstance(symbol_a, str)
        assert isinstance(symbol_b, str)
        idx_a = self.get_grid_id(symbol_a)
        idx_b = self.get_grid_id(symbol_b)
        if idx_a == idx_b:
            return np.arange(self.get_grid(symbol_a).size)
        if idx_a is not None and idx_b is not None:
            ptr = cext.MultiGrid_get_common_index(self._this, idx_a, idx_b)
            size = cext.MultiGrid_get_common_index_size(self._this, idx_a, idx_b)
            return np.ctypeslib.as_array(ptr, shape=(size, 2))
        return None

    def show(self):
        print('\nNumber of Grids: ', self.ngrids)
        print('Object Information: ', self._grids)

    @property
    def ngrids(self):
        return cext.MultiGrid_get_ngrids(self._this)

    @property
    def nspecies(self):
        return cext.MultiGrid_get_nspecies(self._this)
