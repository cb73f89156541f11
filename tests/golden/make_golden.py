"""Writes tests/golden/ref_vectors.npz: input/output vectors of the reference's own
standalone C++ routines, run from oracle/_ref/libnapmo_ref.so (``make -C oracle ref``
builds it from the reference sources where they lie). Only numbers are stored.

Usage: python tests/golden/make_golden.py
"""
import ctypes
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, '..', '..'))

from oracle import ref  # noqa: E402

LEB_ORDERS = [6, 14, 26, 38, 50, 74, 86, 110, 146, 170, 194, 230, 266, 302, 350, 434, 590]


def main():
    L = ref.lib()
    out = {}
    rng = np.random.default_rng(20261202)
    for n in LEB_ORDERS:
        xyz, w = ref.lebedev(n)
        out['leb_%d_xyz' % n] = xyz
        out['leb_%d_w' % n] = w
    # radial transforms
    t = np.arange(100, dtype=np.float64)
    cheb = L.ChebyshevRTransform_new(0.5858151015155881, 100)
    powr = L.PowerRTransform_new(1e-3, 2e1, 100)
    for name, ptr in (('cheb', cheb), ('power', powr)):
        for which in ('radius', 'deriv', 'deriv2', 'deriv3'):
            out['rtf_%s_%s' % (name, which)] = ref.rtransform_all(ptr, t, which)
        r = ref.rtransform_all(ptr, t, 'radius')
        out['rtf_%s_inv' % name] = ref.rtransform_all(ptr, r, 'inv')
    # spline derivatives
    y = rng.normal(size=40)
    out['spline_y'] = y
    out['spline_dt'] = ref.spline_dt(y)
    # build_ode2
    n = 12
    b, a, f = rng.normal(size=2 * n), rng.normal(size=2 * n), rng.normal(size=2 * n)
    bcs = np.array([0.3, np.nan, -0.2, np.nan])
    coeffs, rhs = ref.build_ode2(b, a, f, bcs, n)
    out.update(ode_b=b, ode_a=a, ode_f=f, ode_bcs=bcs, ode_coeffs=coeffs, ode_rhs=rhs)
    # solid harmonics
    vecs = rng.normal(size=(6, 3))
    polys = []
    for v in vecs:
        work = np.zeros(48)
        work[:3] = (v[2], v[0], v[1])
        L.ref_fill_pure_polynomials(work, 6)
        polys.append(work.copy())
    out['pure_vecs'] = vecs
    out['pure_polys'] = np.array(polys)
    # eval_decomposition_grid with potential splines (lmax 3, 20-point Chebyshev grid)
    nrad, lmax = 20, 3
    rtf = L.ChebyshevRTransform_new(1.2, nrad)
    ys = rng.normal(size=((lmax + 1) ** 2, nrad)) * np.exp(-np.arange(nrad) / 5.0)
    dts = rng.normal(size=((lmax + 1) ** 2, nrad)) * 0.1
    keep, sp = [], []
    for s in range((lmax + 1) ** 2):
        l = int(np.sqrt(s))
        yy, dd = np.ascontiguousarray(ys[s]), np.ascontiguousarray(dts[s])
        keep += [yy, dd]
        sp.append(L.CubicSpline_new(yy, dd, L.PotentialExtrapolation_new(l), rtf, nrad))
    center = np.array([0.1, -0.2, 0.3])
    pts = np.concatenate([rng.normal(size=(40, 3)) * 2.0,
                          center + rng.normal(size=(4, 3)) * 1e-6,   # left extrapolation
                          rng.normal(size=(4, 3)) * 3e3])           # right extrapolation
    dec = np.zeros(len(pts))
    L.eval_decomposition_grid(np.array(sp, dtype=ctypes.c_void_p), center, dec,
                              np.ascontiguousarray(pts).reshape(-1), L.Cell_new(None, 0), len(sp), len(pts))
    out.update(dec_y=ys, dec_dt=dts, dec_center=center, dec_points=pts, dec_out=dec,
               dec_radii=np.array([1.2]), dec_lmax=np.array([lmax]))
    # dot_multi_moments (mtype 4, pure moments of the unit vector) per segment
    npt, nseg = 30, 3
    pts2 = rng.normal(size=(npt, 3))
    fvals = rng.normal(size=npt)
    seg = np.array([10, 10, 10], dtype=np.int64)
    mom = np.zeros(nseg * 16)
    L.dot_multi_moments(npt, 1, fvals, np.ascontiguousarray(pts2).reshape(-1), center, 3, 4, seg, mom, 16)
    out.update(mom_points=pts2, mom_f=fvals, mom_out=mom.reshape(nseg, 16))
    # collocation
    shells = [((0.1, 0.2, 0.3), 0, [13.01, 1.962, 0.4446], [0.019685, 0.137977, 0.478148]),
              ((0.1, 0.2, 0.3), 1, [0.727], [1.0]),
              ((-0.5, 0.4, 0.0), 2, [0.8, 2.5], [0.6, 0.4])]
    rb = ref.RefBasis(shells)
    pts3 = rng.normal(size=(25, 3))
    out['gto_points'] = pts3
    out['gto_phi'] = rb.compute(pts3)
    # dot_multi_moments for every mtype: two integrand vectors, uneven segments, output
    # accumulated onto non-zero starting values; then an empty segment, after which the
    # reference keeps every later point in that segment
    rng2 = np.random.default_rng(7)
    npt = 57
    pts4 = rng2.normal(size=(npt, 3)) * 1.5
    data4 = rng2.normal(size=(2, npt))
    center4 = np.array([0.2, -0.1, 0.05])
    counts = {1: lambda l: (l + 1) * (l + 2) * (l + 3) // 6, 2: lambda l: (l + 1) ** 2,
              3: lambda l: l + 1, 4: lambda l: (l + 1) ** 2}
    out.update(mm_points=pts4, mm_data=data4, mm_center=center4)
    for mtype in (1, 2, 3, 4):
        lmax = 5
        nm = counts[mtype](lmax)
        seg = np.array([20, 7, 30], dtype=np.int64)
        init = rng2.normal(size=3 * nm)
        mom = init.copy()
        L.dot_multi_moments(npt, 2, np.ascontiguousarray(data4).reshape(-1),
                            np.ascontiguousarray(pts4).reshape(-1), center4, lmax, mtype, seg, mom, nm)
        out['mm%d_init' % mtype] = init.reshape(3, nm)
        out['mm%d_out' % mtype] = mom.reshape(3, nm)
    seg = np.array([10, 0, 5, 42], dtype=np.int64)
    mom = np.zeros(4 * 16)
    L.dot_multi_moments(npt, 1, np.ascontiguousarray(data4[0]), np.ascontiguousarray(pts4).reshape(-1),
                        center4, 3, 2, seg, mom, 16)
    out['mm_empty_seg'] = seg
    out['mm_empty_out'] = mom.reshape(4, 16)
    np.savez_compressed(os.path.join(HERE, 'ref_vectors.npz'), **out)
    print('wrote', os.path.join(HERE, 'ref_vectors.npz'), 'keys:', len(out))


if __name__ == '__main__':
    main()
