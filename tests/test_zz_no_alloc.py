"""napmo_epc17_2_grid_noalloc (the allocation-free entry point, scratch in a static device
array) against napmo_epc17_2_grid on the same inputs, bit for bit, on two calls in a row:
the first on every common point with one density below 1e-200 (the simplified-form switch
sets the scratch flag), the second on fewer points without it (the flag must be cleared)."""
import ctypes

import numpy as np
import pytest

from tests.test_gpu_glue import E_ATOMS, _species_grids

pytestmark = pytest.mark.gpu


class _Dev(object):
    """Raw device buffer of a host array (any dtype)."""

    def __init__(self, a):
        from napmo_amd.system import cext
        self.a = np.ascontiguousarray(a)
        self.ptr = cext.napmo_device_malloc(max(self.a.nbytes, 8))
        assert self.ptr
        cext.check(cext.napmo_memcpy(self.ptr, self.a.ctypes.data, self.a.nbytes, 1))

    def host(self):
        from napmo_amd.system import cext
        out = np.empty_like(self.a)
        cext.check(cext.napmo_memcpy(out.ctypes.data, self.ptr, out.nbytes, 2))
        return out

    def __del__(self):
        from napmo_amd.system import cext
        if getattr(self, 'ptr', None):
            cext.napmo_device_free(self.ptr)
            self.ptr = None


def _existing(ge, gp, idx, rho_e, rho_p, va, vb):
    from napmo_amd.system import cext
    va, vb = va.copy(), vb.copy()
    e = np.zeros(1)
    ii = np.ascontiguousarray(idx, dtype=np.int32)
    cext.check(cext.napmo_epc17_2_grid(ge._this, gp._this, ii.ctypes.data, len(ii), rho_e.ctypes.data,
                                       rho_p.ctypes.data, va.ctypes.data, vb.ctypes.data, e.ctypes.data))
    return va, vb, e[0]


def _noalloc(ge, gp, idx, rho_e, rho_p, va, vb):
    from napmo_amd.system import cext
    d = [_Dev(np.ascontiguousarray(idx, dtype=np.int32)), _Dev(rho_e), _Dev(rho_p), _Dev(va), _Dev(vb),
         _Dev(np.zeros(1))]
    cext.check(cext.napmo_epc17_2_grid_noalloc(ge._this, gp._this, d[0].ptr, len(idx), d[1].ptr, d[2].ptr,
                                               d[3].ptr, d[4].ptr, d[5].ptr))
    cext.check(cext.napmo_grid_sync(gp._this))
    return d[3].host(), d[4].host(), d[5].host()[0]


def test_noalloc_entry_point_matches_the_allocating_one_on_two_calls():
    import napmo_amd
    ge, gp = _species_grids()
    mg = napmo_amd.MultiGrid(2)
    mg.add_grid(ge)
    mg.add_grid(gp)
    idx = np.array(mg.get_common_points('e-', 'H+'))
    assert len(idx) > 100
    pe, pp = ge.points, gp.points
    rho_e = sum(np.exp(-((pe - np.array(c)) ** 2).sum(1)) for _, c in E_ATOMS)
    rho_p = 30.0 * np.exp(-8.0 * ((pp - np.array(E_ATOMS[1][1])) ** 2).sum(1))
    va, vb = np.full(ge.size, 0.25), np.full(gp.size, -0.5)
    tiny = rho_e.copy()
    tiny[idx[3, 0]] = 0.0   # below 1e-200: the whole call takes the simplified form
    for rows, re in ((idx, tiny), (idx[: len(idx) // 2], rho_e)):
        ref = _existing(ge, gp, rows, re, rho_p, va, vb)
        got = _noalloc(ge, gp, rows, re, rho_p, va, vb)
        print('npair %d: energy %.17g vs %.17g' % (len(rows), got[2], ref[2]))
        assert np.array_equal(got[0], ref[0]) and np.array_equal(got[1], ref[1])
        assert got[2] == ref[2]
