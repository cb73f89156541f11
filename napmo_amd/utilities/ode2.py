"""Finite-element second-order ODE solver (napmo/utilities/ode2.py:35-118).

The Galerkin system is assembled by the library's ``build_ode2`` and solved with a
dense FP64 LU (numpy), then returned as a CubicSpline like the reference.
"""
import numpy as np

from napmo_amd.grids.cubic_spline import CubicSpline
from napmo_amd.system import cext


def solve_ode2(b, a, f, bcs, extrapolation=None, j1=None, j2=None, j3=None):
    def merge(y, d):
        return np.ascontiguousarray(np.array([y, d]).T.ravel())

    rtf = b.rtransform
    if rtf.to_string() != a.rtransform.to_string():
        raise ValueError('The RTransform objects of b and a do not match.')
    if rtf.to_string() != f.rtransform.to_string():
        raise ValueError('The RTransform objects of b and f do not match.')
    j1 = rtf.deriv_all() if j1 is None else j1
    j2 = rtf.deriv2_all() if j2 is None else j2
    j3 = rtf.deriv3_all() if j3 is None else j3
    j1sq = j1 * j1
    by_new = j1 * b.y - j2 / j1
    bd_new = j2 * b.y + j1sq * b.dx + (j2 * j2 - j1 * j3) / j1sq
    ay_new = a.y * j1sq
    ad_new = (a.dx * j1sq + 2 * a.y * j2) * j1
    fy_new = f.y * j1sq
    fd_new = (f.dx * j1sq + 2 * f.y * j2) * j1
    new_bcs = (bcs[0], None if bcs[1] is None else bcs[1] * j1[0],
               bcs[2], None if bcs[3] is None else bcs[3] * j1[-1])
    npoint = by_new.shape[0]
    c_bcs = np.array([np.nan if v is None else v for v in new_bcs], dtype=np.float64)
    coeffs = np.zeros((2 * npoint, 2 * npoint), dtype=np.float64)
    rhs = np.zeros(2 * npoint, dtype=np.float64)
    cext.build_ode2(merge(by_new, bd_new), merge(ay_new, ad_new), merge(fy_new, fd_new), c_bcs,
                    coeffs, rhs, npoint)
    solution = np.linalg.solve(coeffs, rhs)
    uy = solution[::2].copy()
    ud = solution[1::2] / j1
    return CubicSpline(uy, ud, rtf, extrapolation)
