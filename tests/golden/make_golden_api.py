"""Writes tests/golden/ref_vectors_api.npz: input/output vectors of the reference's spline,
extrapolation and Gaussian routines behind the drop-in symbols (CubicSpline_eval_deriv,
Extrapolation_*, compute_cubic_spline_int_weights, PrimitiveGaussian_* and
ContractedGaussian_* compute / overlap, BasisSet_get_max_*), run from
oracle/_ref/libnapmo_ref.so (``make -C oracle ref`` builds it from the reference sources
where they lie). Only numbers are stored.

Usage: python tests/golden/make_golden_api.py
"""
import ctypes
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, '..', '..'))

from oracle import ref  # noqa: E402

d1 = ref.d1
vp, ci, cd = ctypes.c_void_p, ctypes.c_int, ctypes.c_double


def bind(L):
    sig = {
        'CubicSpline_eval_deriv': (None, [vp, d1, d1, ci]),
        'ZeroExtrapolation_new': (vp, []),
        'CuspExtrapolation_new': (vp, []),
        'PowerExtrapolation_new': (vp, [cd]),
        'PowerExtrapolation_get_power': (cd, [vp]),
        'Extrapolation_eval_left': (cd, [vp, cd]),
        'Extrapolation_eval_right': (cd, [vp, cd]),
        'Extrapolation_deriv_left': (cd, [vp, cd]),
        'Extrapolation_deriv_right': (cd, [vp, cd]),
        'Extrapolation_has_tail': (ctypes.c_bool, [vp]),
        'compute_cubic_spline_int_weights': (None, [d1, ci]),
        'PrimitiveGaussian_compute': (None, [vp, d1, d1, ci]),
        'PrimitiveGaussian_overlap': (cd, [vp, vp]),
        'PrimitiveGaussian_get_norma': (cd, [vp]),
        'ContractedGaussian_compute': (None, [vp, d1, d1, ci]),
        'ContractedGaussian_overlap': (cd, [vp, vp]),
        'ContractedGaussian_get_nprim': (ci, [vp]),
        'ContractedGaussian_get_norma': (cd, [vp]),
        'BasisSet_get_max_l': (ci, [vp]),
        'BasisSet_get_max_nprim': (ci, [vp]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype, fn.argtypes = res, args


def main():
    L = ref.lib()
    bind(L)
    out = {}
    rng = np.random.default_rng(20261203)
    # splines on a 30-point Chebyshev grid with every extrapolation kind
    n = 30
    rtf = L.ChebyshevRTransform_new(0.8, n)
    r = ref.rtransform_all(rtf, np.arange(n, dtype=np.float64), 'radius')
    y = np.ascontiguousarray(np.exp(-0.7 * r) * (1.0 + 0.3 * np.sin(3.0 * r)) + 0.05 / (1.0 + r))
    dt = ref.spline_dt(y)
    x = np.concatenate([r[0] * np.array([0.2, 0.5, 0.9]),
                        np.sort(rng.uniform(r[0], r[-1], 40)), r, r[-1] * np.array([1.1, 2.0, 7.5])])
    x = np.ascontiguousarray(x)
    out.update(sp_r=r, sp_y=y, sp_dt=dt, sp_x=x)
    kinds = {'zero': lambda: L.ZeroExtrapolation_new(), 'cusp': lambda: L.CuspExtrapolation_new(),
             'power': lambda: L.PowerExtrapolation_new(-2.5), 'pot0': lambda: L.PotentialExtrapolation_new(0),
             'pot2': lambda: L.PotentialExtrapolation_new(2)}
    probes = np.array([r[0] * 0.3, r[0] * 0.7, r[-1] * 1.5, r[-1] * 4.0])
    out['ex_probes'] = probes
    for name, make in kinds.items():
        ex = make()
        cs = L.CubicSpline_new(y, dt, ex, rtf, n)
        val, der = np.zeros(len(x)), np.zeros(len(x))
        L.CubicSpline_eval(cs, x, val, len(x))
        L.CubicSpline_eval_deriv(cs, x, der, len(x))
        out['sp_%s_val' % name] = val
        out['sp_%s_der' % name] = der
        out['ex_%s' % name] = np.array([[L.Extrapolation_eval_left(ex, p), L.Extrapolation_eval_right(ex, p),
                                         L.Extrapolation_deriv_left(ex, p), L.Extrapolation_deriv_right(ex, p)]
                                        for p in probes])
        out['ex_%s_tail' % name] = np.array([bool(L.Extrapolation_has_tail(ex))])
        if name == 'power':
            out['ex_power_power'] = np.array([L.PowerExtrapolation_get_power(ex)])
    # integration weights
    for m in (2, 10, 25):
        w = np.zeros(m)
        L.compute_cubic_spline_int_weights(w, m)
        out['intw_%d' % m] = w
    # primitives and contractions
    pts = np.ascontiguousarray(rng.normal(size=(20, 3)) * 1.3)
    out['g_points'] = pts
    prims = [((1, 0, 2), (0.1, -0.3, 0.2), 0.9, 0.7), ((0, 0, 0), (0.0, 0.0, 0.0), 3.2, 0.4),
             ((2, 1, 0), (-0.6, 0.5, 0.1), 0.35, 1.1), ((0, 1, 0), (0.2, 0.2, -0.4), 12.0, 0.25)]
    out['g_prims_l'] = np.array([p[0] for p in prims], dtype=np.int32)
    out['g_prims_o'] = np.array([p[1] for p in prims])
    out['g_prims_zc'] = np.array([(p[2], p[3]) for p in prims])
    ptr = []
    vals = []
    for l, o, z, c in prims:
        p = L.PrimitiveGaussian_new(np.array(l, dtype=np.int32), np.array(o), z, c)
        ptr.append(p)
        v = np.zeros(len(pts))
        L.PrimitiveGaussian_compute(p, pts.reshape(-1), v, len(pts))
        vals.append(v)
    out['g_prim_val'] = np.array(vals)
    out['g_prim_norma'] = np.array([L.PrimitiveGaussian_get_norma(p) for p in ptr])
    out['g_prim_ovl'] = np.array([[L.PrimitiveGaussian_overlap(a, b) for b in ptr] for a in ptr])
    # two contractions of same-l primitives (the reference takes l and origin from the first)
    conts = []
    groups = [[((1, 1, 0), (0.3, 0.1, -0.2), z, c) for z, c in ((5.0, 0.2), (1.2, 0.5), (0.3, 0.6))],
              [((0, 0, 2), (-0.4, 0.0, 0.5), z, c) for z, c in ((2.0, 0.8), (0.5, 0.3))]]
    keep = []
    for gi, grp in enumerate(groups):
        ps = [L.PrimitiveGaussian_new(np.array(l, dtype=np.int32), np.array(o), z, c) for l, o, z, c in grp]
        arr = np.array(ps, dtype=ctypes.c_void_p)
        keep.append(arr)
        cg = L.ContractedGaussian_new(arr, len(ps))
        conts.append(cg)
        v = np.zeros(len(pts))
        L.ContractedGaussian_compute(cg, pts.reshape(-1), v, len(pts))
        out['g_cont%d_val' % gi] = v
        out['g_cont%d_def' % gi] = np.array([[l[0], l[1], l[2], o[0], o[1], o[2], z, c] for l, o, z, c in grp])
        out['g_cont%d_nprim' % gi] = np.array([L.ContractedGaussian_get_nprim(cg)])
        out['g_cont%d_norma' % gi] = np.array([L.ContractedGaussian_get_norma(cg)])
    out['g_cont_ovl'] = np.array([[L.ContractedGaussian_overlap(a, b) for b in conts] for a in conts])
    rb = ref.RefBasis([((0.1, 0.2, 0.3), 0, [13.01, 1.962, 0.4446], [0.019685, 0.137977, 0.478148]),
                       ((0.1, 0.2, 0.3), 1, [0.727], [1.0]),
                       ((-0.5, 0.4, 0.0), 2, [0.8, 2.5], [0.6, 0.4])])
    out['basis_max_l'] = np.array([L.BasisSet_get_max_l(rb.ptr)])
    out['basis_max_nprim'] = np.array([L.BasisSet_get_max_nprim(rb.ptr)])
    np.savez_compressed(os.path.join(HERE, 'ref_vectors_api.npz'), **out)
    print('wrote', os.path.join(HERE, 'ref_vectors_api.npz'), 'keys:', len(out))


if __name__ == '__main__':
    main()
