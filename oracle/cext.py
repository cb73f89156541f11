"""ctypes binding of oracle/liboracle.so (C/OpenMP restatement in napmo_oracle.c).
TEST INFRASTRUCTURE ONLY (parity checker and CPU baseline)."""
import ctypes
import os
import subprocess

import numpy as np

from oracle import core

_HERE = os.path.dirname(os.path.abspath(__file__))
_PATH = os.path.join(_HERE, 'liboracle.so')
_lib = None
_TABS = None

d1 = np.ctypeslib.ndpointer(dtype=np.float64, flags='C_CONTIGUOUS')
i1 = np.ctypeslib.ndpointer(dtype=np.int32, flags='C_CONTIGUOUS')
l1 = np.ctypeslib.ndpointer(dtype=np.int64, flags='C_CONTIGUOUS')
C = ctypes


def build():
    subprocess.check_call(['make', '-s', '-C', _HERE, 'liboracle.so'])


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_PATH):
            build()
        L = ctypes.CDLL(_PATH)
        L.oracle_build_ode2.argtypes = [d1, d1, d1, d1, d1, d1, C.c_long, d1, d1]
        L.oracle_becke_weights.argtypes = [d1, C.c_long, d1, d1, C.c_int, l1, d1]
        L.oracle_basis_compute.argtypes = [d1, C.c_long, C.c_int, d1, i1, i1, i1, d1, d1, d1, d1]
        L.oracle_eval_decomposition.argtypes = [d1, d1, d1, d1, C.c_int, C.c_int, C.c_int, C.c_double,
                                                C.c_double, C.c_double, C.c_double, d1, d1, C.c_long, d1]
        L.oracle_coupling.argtypes = [d1, d1, d1, d1, C.c_long, C.c_int, d1]
        L.oracle_density.argtypes = [d1, d1, C.c_long, C.c_int, d1]
        L.oracle_set_threads.argtypes = [C.c_int]
        L.oracle_set_threads.restype = None
        L.oracle_max_threads.restype = C.c_int
        for fn in ('oracle_build_ode2', 'oracle_becke_weights', 'oracle_basis_compute',
                   'oracle_eval_decomposition', 'oracle_coupling', 'oracle_density'):
            getattr(L, fn).restype = None
        _lib = L
    return _lib


def hermite_tables():
    """Per-segment integral tables for oracle_build_ode2 (3 bits per factor:
    relative node, kind, derivative), from the exact integrals of oracle/core.py."""
    global _TABS
    if _TABS is None:
        t2 = np.zeros(64)
        t3 = np.zeros(512)

        def f(bits):
            return ((bits >> 2) & 1, (bits >> 1) & 1, bool(bits & 1))
        for a in range(8):
            for b in range(8):
                t2[(a << 3) | b] = core._seg_overlap([f(a), f(b)], 0)
                for c in range(8):
                    t3[(a << 6) | (b << 3) | c] = core._seg_overlap([f(a), f(b), f(c)], 0)
        _TABS = (t2, t3)
    return _TABS


def build_ode2(b, a, f, bcs, npoint):
    t2, t3 = hermite_tables()
    coeffs = np.zeros(4 * npoint * npoint)
    rhs = np.zeros(2 * npoint)
    lib().oracle_build_ode2(np.ascontiguousarray(b, dtype=np.float64), np.ascontiguousarray(a, dtype=np.float64),
                            np.ascontiguousarray(f, dtype=np.float64), np.asarray(bcs, dtype=np.float64),
                            coeffs, rhs, npoint, t2, t3)
    return coeffs.reshape(2 * npoint, 2 * npoint), rhs


def becke_weights(points, origin, radii, offsets):
    out = np.zeros(points.shape[0])
    lib().oracle_becke_weights(np.ascontiguousarray(points).reshape(-1), points.shape[0],
                               np.ascontiguousarray(origin, dtype=np.float64).reshape(-1),
                               np.ascontiguousarray(radii, dtype=np.float64), len(radii),
                               np.asarray(offsets, dtype=np.int64), out)
    return out


def flatten_basis(basis):
    fo, fl, fnp, ff, fn, pz, pc = [], [], [], [], [], [], []
    for o, lxyz, zetas, cn, norm in basis.functions:
        fo.extend(o)
        fl.extend(lxyz)
        fnp.append(len(zetas))
        ff.append(len(pz))
        fn.append(norm)
        pz.extend(zetas)
        pc.extend(cn)
    return (np.array(fo, dtype=np.float64), np.array(fl, dtype=np.int32), np.array(fnp, dtype=np.int32),
            np.array(ff, dtype=np.int32), np.array(fn, dtype=np.float64), np.array(pz, dtype=np.float64),
            np.array(pc, dtype=np.float64))


def basis_compute(basis, points):
    out = np.zeros((points.shape[0], basis.nbasis))
    lib().oracle_basis_compute(np.ascontiguousarray(points).reshape(-1), points.shape[0], basis.nbasis,
                               *flatten_basis(basis), out.reshape(-1))
    return out


def rtf_params(rtf):
    if rtf.kind == 'chebyshev':
        return 0, rtf.radii, 0.0
    if rtf.kind == 'power':
        return 1, rtf.rmin, rtf.power
    return 2, 0.0, 0.0


def eval_decomposition(splines, center, points, out):
    n = splines[0].n
    y = np.ascontiguousarray(np.array([s.y for s in splines]))
    dt = np.ascontiguousarray(np.array([s.dt for s in splines]))
    al = np.array([s.extrapolation.amp_left for s in splines])
    ar = np.array([s.extrapolation.amp_right for s in splines])
    k, p0, p1 = rtf_params(splines[0].rtf)
    lib().oracle_eval_decomposition(y.reshape(-1), dt.reshape(-1), al, ar, n, len(splines), k, p0, p1,
                                    splines[0].first_x, splines[0].last_x,
                                    np.asarray(center, dtype=np.float64), np.ascontiguousarray(points).reshape(-1),
                                    points.shape[0], out)
    return out


def coupling(phi_fm, v, bw, w):
    """phi_fm: function-major [n, N]."""
    n, N = phi_fm.shape
    out = np.zeros(n * n)
    lib().oracle_coupling(np.ascontiguousarray(phi_fm).reshape(-1), np.ascontiguousarray(v),
                          np.ascontiguousarray(bw), np.ascontiguousarray(w), N, n, out)
    return out.reshape(n, n)


def density(phi, P):
    out = np.zeros(phi.shape[0])
    lib().oracle_density(np.ascontiguousarray(phi).reshape(-1), np.ascontiguousarray(P).reshape(-1),
                         phi.shape[0], phi.shape[1], out)
    return out


def set_threads(n):
    lib().oracle_set_threads(int(n))


def max_threads():
    return int(lib().oracle_max_threads())
