"""Registry of several species' Becke grids and their shared points.

Public surface of napmo/grids/multi_grid.py:13-146 (MultiGrid(nspecies), add_grid,
get_grid, get_grid_id, get_common_points, show, ngrids, nspecies). The common-point index
is computed by the library (src/grid/multi_grid.cpp:37-86 in the reference): here the
smaller grid's coordinates are hashed by bit pattern, and the rows come out in the
reference's order.

Lookup rules kept from the reference API:
  * get_grid(symbol, gid): a registered symbol wins; otherwise the grid registered under
    id ``gid``; otherwise None.
  * get_grid_id(symbol): the library id of that symbol's grid, or None.
  * get_common_points(a, b): for one and the same grid, arange(size) (every point is
    common); for two registered grids, a read-only [n x 2] int view into the library's
    index (column 0 indexes the larger grid, ``a`` on a tie); None if either is unknown.
"""
import numpy as np

from napmo_amd.system import cext


class MultiGrid(object):

    def __init__(self, nspecies):
        assert isinstance(nspecies, int)
        self._ids = {}      # species symbol -> library grid id
        self._by_id = {}    # library grid id -> BeckeGrid
        self._this = cext.MultiGrid_new(nspecies)

    def __del__(self):
        if getattr(self, '_this', None):
            cext.MultiGrid_del(self._this)
            self._this = None

    def add_grid(self, grid):
        from napmo_amd.grids.becke import BeckeGrid
        assert isinstance(grid, BeckeGrid)
        gid = int(cext.MultiGrid_add_grid(self._this, grid._this))
        old = self._ids.get(grid._symbol)
        if old is not None and old != gid:
            self._by_id.pop(old, None)
        self._ids[grid._symbol] = gid
        self._by_id[gid] = grid

    def get_grid(self, symbol='', gid=-1):
        assert isinstance(symbol, str)
        assert isinstance(gid, int)
        if symbol in self._ids:
            return self._by_id[self._ids[symbol]]
        return self._by_id.get(gid)

    def get_grid_id(self, symbol):
        assert isinstance(symbol, str)
        return self._ids.get(symbol)

    def get_common_points(self, symbol_a, symbol_b):
        assert isinstance(symbol_a, str)
        assert isinstance(symbol_b, str)
        ga, gb = self._ids.get(symbol_a), self._ids.get(symbol_b)
        if ga == gb:
            if ga is None:
                return None
            return np.arange(self._by_id[ga].size)
        if ga is None or gb is None:
            return None
        ptr = cext.MultiGrid_get_common_index(self._this, ga, gb)  # computes the pair's index
        nrow = int(cext.MultiGrid_get_common_index_size(self._this, ga, gb))
        if nrow <= 0 or not ptr:
            return np.zeros((0, 2), dtype=np.int32)
        return np.ctypeslib.as_array(ptr, shape=(nrow, 2))

    def show(self):
        print('\nMultiGrid: %d grid(s), %d species' % (self.ngrids, self.nspecies))
        for sym, gid in sorted(self._ids.items(), key=lambda kv: kv[1]):
            print('  id %d  %-8s %d points' % (gid, sym, self._by_id[gid].size))

    @property
    def ngrids(self):
        return cext.MultiGrid_get_ngrids(self._this)

    @property
    def nspecies(self):
        return cext.MultiGrid_get_nspecies(self._this)
