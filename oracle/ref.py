"""ctypes binding of oracle/_ref/libnapmo_ref.so: the reference's own standalone C++
sources compiled where they lie (``make -C oracle ref``). TEST INFRASTRUCTURE ONLY:
used by tests to pin the restatement in ``oracle/core.py``; absent on machines
without the reference sources (tests that need it skip)."""
import ctypes
import os

import numpy as np

_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), '_ref', 'libnapmo_ref.so')
_lib = None

d1 = np.ctypeslib.ndpointer(dtype=np.float64, flags='C_CONTIGUOUS')
i1 = np.ctypeslib.ndpointer(dtype=np.int32, flags='C_CONTIGUOUS')
l1 = np.ctypeslib.ndpointer(dtype=np.int64, flags='C_CONTIGUOUS')
ptrs = np.ctypeslib.ndpointer(dtype=ctypes.c_void_p, flags='C_CONTIGUOUS')
vp = ctypes.c_void_p


def available():
    return os.path.exists(_PATH)


def lib():
    global _lib
    if _lib is None:
        L = ctypes.CDLL(_PATH)
        sig = {
            'ref_ld_by_order': (None, [ctypes.c_int, d1, d1, d1, d1]),
            'ref_fill_pure_polynomials': (ctypes.c_long, [d1, ctypes.c_long]),
            'ChebyshevRTransform_new': (vp, [ctypes.c_double, ctypes.c_int]),
            'PowerRTransform_new': (vp, [ctypes.c_double, ctypes.c_double, ctypes.c_int]),
            'RTransform_radius_array': (None, [vp, d1, d1, ctypes.c_int]),
            'RTransform_deriv_array': (None, [vp, d1, d1, ctypes.c_int]),
            'RTransform_deriv2_array': (None, [vp, d1, d1, ctypes.c_int]),
            'RTransform_deriv3_array': (None, [vp, d1, d1, ctypes.c_int]),
            'RTransform_inv_array': (None, [vp, d1, d1, ctypes.c_int]),
            'solve_cubic_spline_system': (None, [d1, d1, ctypes.c_int]),
            'build_ode2': (None, [d1, d1, d1, d1, d1, d1, ctypes.c_long]),
            'PotentialExtrapolation_new': (vp, [ctypes.c_int]),
            'CubicSpline_new': (vp, [d1, d1, vp, vp, ctypes.c_int]),
            'CubicSpline_eval': (None, [vp, d1, d1, ctypes.c_int]),
            'Cell_new': (vp, [ctypes.c_void_p, ctypes.c_int]),
            'eval_decomposition_grid': (None, [ptrs, d1, d1, d1, vp, ctypes.c_long, ctypes.c_long]),
            'dot_multi_moments': (None, [ctypes.c_long, ctypes.c_long, d1, d1, d1, ctypes.c_long,
                                         ctypes.c_long, l1, d1, ctypes.c_long]),
            'PrimitiveGaussian_new': (vp, [i1, d1, ctypes.c_double, ctypes.c_double]),
            'ContractedGaussian_new': (vp, [ptrs, ctypes.c_int]),
            'BasisSet_new': (vp, [ptrs, ctypes.c_int]),
            'BasisSet_compute': (None, [vp, d1, d1, ctypes.c_int]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def lebedev(n):
    out = [np.zeros(n) for _ in range(4)]
    lib().ref_ld_by_order(n, *out)
    return np.stack(out[:3], axis=1), out[3]


def rtransform_all(rtf_ptr, t, which):
    out = np.zeros(len(t))
    getattr(lib(), 'RTransform_%s_array' % which)(rtf_ptr, np.ascontiguousarray(t), out, len(t))
    return out


def build_ode2(b, a, f, bcs, npoint):
    coeffs = np.zeros((2 * npoint, 2 * npoint))
    rhs = np.zeros(2 * npoint)
    lib().build_ode2(np.ascontiguousarray(b), np.ascontiguousarray(a), np.ascontiguousarray(f),
                     np.asarray(bcs, dtype=np.float64), coeffs.reshape(-1), rhs, npoint)
    return coeffs, rhs


def spline_dt(y):
    y = np.ascontiguousarray(y, dtype=np.float64)
    dt = np.zeros_like(y)
    lib().solve_cubic_spline_system(y, dt, len(y))
    return dt


class RefBasis:
    """Reference BasisSet built from (origin, l, zetas, coeffs) shells."""

    def __init__(self, shells):
        from oracle.core import cartesian_powers
        self._keep = []
        conts = []
        for origin, l, zetas, coeffs in shells:
            for lxyz in cartesian_powers(l):
                ll = np.array(lxyz, dtype=np.int32)
                o = np.asarray(origin, dtype=np.float64)
                prims = [lib().PrimitiveGaussian_new(ll, o, float(z), float(c)) for z, c in zip(zetas, coeffs)]
                arr = np.array(prims, dtype=ctypes.c_void_p)
                self._keep.append(arr)
                conts.append(lib().ContractedGaussian_new(arr, len(prims)))
        self._conts = np.array(conts, dtype=ctypes.c_void_p)
        self.nbasis = len(conts)
        self.ptr = lib().BasisSet_new(self._conts, self.nbasis)

    def compute(self, points):
        points = np.ascontiguousarray(points, dtype=np.float64)
        out = np.zeros((points.shape[0], self.nbasis))
        lib().BasisSet_compute(self.ptr, points.reshape(-1), out.reshape(-1), points.shape[0])
        return out
