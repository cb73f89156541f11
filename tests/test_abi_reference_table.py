"""The reference's ctypes declarations bind libnapmo_gpu.so unchanged, and the host-side
symbols behind them reproduce the reference's own routines (golden vectors from
oracle/_ref in tests/golden/ref_vectors_api.npz). Symbols that need the GPU are called
through the same binding in tests/test_gpu_reference_table.py."""
import ctypes
import os

import numpy as np
import pytest

from tests import reftable, systems

GOLD = np.load(os.path.join(systems.ROOT, 'tests', 'golden', 'ref_vectors_api.npz'))
HOST_SYMBOLS = None  # filled by test_host_symbols


def test_every_grid_path_symbol_is_exported():
    lib = ctypes.CDLL(reftable.LIBPATH)
    names = reftable.grid_path_symbols()
    assert len(names) == 121
    assert [n for n in names if not hasattr(lib, n)] == []


def _spline_check(L, kind_new, name):
    g = GOLD
    n = len(g['sp_r'])
    rtf = L.ChebyshevRTransform_new(0.8, n)
    y, dt = np.ascontiguousarray(g['sp_y']), np.zeros(n)
    L.solve_cubic_spline_system(y, dt, n)
    np.testing.assert_allclose(dt, g['sp_dt'], rtol=0, atol=1e-14)
    ex = kind_new()
    cs = L.CubicSpline_new(y, dt, ex, rtf, n)
    assert L.CubicSpline_get_extrapolation(cs) == ex
    assert L.CubicSpline_get_rtransform(cs) == rtf
    assert L.CubicSpline_get_first_x(cs) == g['sp_r'][0] and L.CubicSpline_get_last_x(cs) == g['sp_r'][-1]
    x = np.ascontiguousarray(g['sp_x'])
    val, der = np.zeros(len(x)), np.zeros(len(x))
    L.CubicSpline_eval(cs, x, val, len(x))
    L.CubicSpline_eval_deriv(cs, x, der, len(x))
    np.testing.assert_allclose(val, g['sp_%s_val' % name], rtol=1e-14, atol=1e-16)
    np.testing.assert_allclose(der, g['sp_%s_der' % name], rtol=1e-12, atol=1e-14)
    got = np.array([[L.Extrapolation_eval_left(ex, p), L.Extrapolation_eval_right(ex, p),
                     L.Extrapolation_deriv_left(ex, p), L.Extrapolation_deriv_right(ex, p)]
                    for p in g['ex_probes']])
    np.testing.assert_allclose(got, g['ex_%s' % name], rtol=1e-14, atol=0)
    assert bool(L.Extrapolation_has_tail(ex)) == bool(g['ex_%s_tail' % name][0])
    L.Extrapolation_prepare(ex, cs)  # re-preparing from the same spline changes nothing
    np.testing.assert_allclose(L.Extrapolation_eval_right(ex, g['ex_probes'][2]), g['ex_%s' % name][2, 1],
                               rtol=1e-14)
    L.CubicSpline_del(cs)
    return ex, rtf


def test_host_symbols_through_the_reference_binding():
    L = reftable.Bound()
    g = GOLD
    # splines and every extrapolation kind
    kinds = {'zero': L.ZeroExtrapolation_new, 'cusp': L.CuspExtrapolation_new,
             'power': lambda: L.PowerExtrapolation_new(-2.5), 'pot0': lambda: L.PotentialExtrapolation_new(0),
             'pot2': lambda: L.PotentialExtrapolation_new(2)}
    for name, make in kinds.items():
        ex, rtf = _spline_check(L, make, name)
        if name == 'power':
            assert L.PowerExtrapolation_get_power(ex) == g['ex_power_power'][0]
        L.Extrapolation_del(ex)
        L.RTransform_del(rtf)
    for m in (2, 10, 25):
        w = np.zeros(m)
        L.compute_cubic_spline_int_weights(w, m)
        np.testing.assert_allclose(w, g['intw_%d' % m], rtol=1e-14, atol=1e-16)
    # primitives
    pts = np.ascontiguousarray(g['g_points'])
    prims = []
    for l, o, (z, c), ref_val, ref_norm in zip(g['g_prims_l'], g['g_prims_o'], g['g_prims_zc'],
                                               g['g_prim_val'], g['g_prim_norma']):
        p = L.PrimitiveGaussian_new(np.ascontiguousarray(l), np.ascontiguousarray(o), z, c)
        prims.append(p)
        v = np.zeros(len(pts))
        L.PrimitiveGaussian_compute(p, pts, v, len(pts))
        np.testing.assert_allclose(v, ref_val, rtol=1e-14, atol=1e-300)
        assert abs(L.PrimitiveGaussian_get_norma(p) - ref_norm) <= 1e-14 * ref_norm
        ll, oo = np.zeros(3, dtype=np.int32), np.zeros(3)
        L.PrimitiveGaussian_get_l(p, ll)
        L.PrimitiveGaussian_get_origin(p, oo)
        assert (ll == l).all() and (oo == o).all()
        assert L.PrimitiveGaussian_get_zeta(p) == z and L.PrimitiveGaussian_get_coeff(p) == c
    ovl = np.array([[L.PrimitiveGaussian_overlap(a, b) for b in prims] for a in prims])
    np.testing.assert_allclose(ovl, g['g_prim_ovl'], rtol=1e-13, atol=1e-15)
    # contractions
    conts, keep = [], []
    for gi in range(2):
        d = g['g_cont%d_def' % gi]
        ps = [L.PrimitiveGaussian_new(np.ascontiguousarray(r[:3].astype(np.int32)), np.ascontiguousarray(r[3:6]),
                                      r[6], r[7]) for r in d]
        arr = np.array(ps, dtype=ctypes.c_void_p)
        keep.append(arr)
        cg = L.ContractedGaussian_new(arr, len(ps))
        conts.append(cg)
        v = np.zeros(len(pts))
        L.ContractedGaussian_compute(cg, pts, v, len(pts))
        np.testing.assert_allclose(v, g['g_cont%d_val' % gi], rtol=1e-13, atol=1e-300)
        assert L.ContractedGaussian_get_nprim(cg) == g['g_cont%d_nprim' % gi][0]
        assert abs(L.ContractedGaussian_get_norma(cg) - g['g_cont%d_norma' % gi][0]) <= 1e-13
        ll, oo = np.zeros(3, dtype=np.int32), np.zeros(3)
        L.ContractedGaussian_get_l(cg, ll)
        L.ContractedGaussian_get_origin(cg, oo)
        assert (ll == d[0, :3]).all() and (oo == d[0, 3:6]).all()
    ovl = np.array([[L.ContractedGaussian_overlap(a, b) for b in conts] for a in conts])
    np.testing.assert_allclose(ovl, g['g_cont_ovl'], rtol=1e-13, atol=1e-15)
    assert abs(ovl[0, 0] - 1.0) < 1e-13  # normalised
    # a basis set of the collocation test's shells: max l and primitive count
    from napmo_amd.gto.basis_set import cartesian_powers
    shells = [((0.1, 0.2, 0.3), 0, [13.01, 1.962, 0.4446], [0.019685, 0.137977, 0.478148]),
              ((0.1, 0.2, 0.3), 1, [0.727], [1.0]), ((-0.5, 0.4, 0.0), 2, [0.8, 2.5], [0.6, 0.4])]
    cl = []
    for o, l, zs, cs in shells:
        for lxyz in cartesian_powers(l):
            arr = np.array([L.PrimitiveGaussian_new(np.array(lxyz, dtype=np.int32), np.array(o), z, c)
                            for z, c in zip(zs, cs)], dtype=ctypes.c_void_p)
            keep.append(arr)
            cl.append(L.ContractedGaussian_new(arr, len(arr)))
    conts_arr = np.array(cl, dtype=ctypes.c_void_p)
    b = L.BasisSet_new(conts_arr, len(cl))
    assert L.BasisSet_get_nbasis(b) == 10
    assert L.BasisSet_get_max_l(b) == g['basis_max_l'][0]
    assert L.BasisSet_get_max_nprim(b) == g['basis_max_nprim'][0]
    e = L.BasisSet_new_empty()
    L.BasisSet_update(e, b)
    assert L.BasisSet_get_nbasis(e) == 10 and L.BasisSet_get_max_l(e) == 2
    # radial transforms, grids and their reductions
    rt = L.ChebyshevRTransform_new(0.5858151015155881, 50)
    assert L.RTransform_get_npoint(rt) == 50 and L.ChebyshevRTransform_get_radii(rt) == 0.5858151015155881
    t = np.arange(50, dtype=np.float64)
    rr = np.zeros(50)
    L.RTransform_radius_array(rt, t, rr, 50)
    for nm in ('deriv', 'deriv2', 'deriv3', 'inv'):
        out = np.zeros(50)
        getattr(L, 'RTransform_%s_array' % nm)(rt, rr if nm == 'inv' else t, out, 50)
        assert abs(out[7] - getattr(L, 'RTransform_' + nm)(rt, rr[7] if nm == 'inv' else 7.0)) < 1e-15 * max(1, abs(out[7]))
    assert abs(L.RTransform_radius(rt, 7.0) - rr[7]) == 0
    pr = L.PowerRTransform_new(1e-3, 20.0, 40)
    assert L.PowerRTransform_get_rmin(pr) == 1e-3 and L.PowerRTransform_get_rmax(pr) == 20.0
    assert L.PowerRTransform_get_power(pr) > 0
    ir = L.IdentityRTransform_new(10)
    assert L.RTransform_radius(ir, 3.0) == 3.0
    rad = L.RadialGrid_new(rt, 0.5858151015155881)
    assert L.RadialGrid_get_size(rad) == 50 and L.RadialGrid_get_radii(rad) == 0.5858151015155881
    rp = np.ctypeslib.as_array(L.RadialGrid_get_points(rad), shape=(50,))
    rw = np.ctypeslib.as_array(L.RadialGrid_get_weights(rad), shape=(50,))
    np.testing.assert_array_equal(rp, rr)
    # closed form (napmo/grids/test/test_radial.py): 4 pi int r^2 exp(-r^2/2) dr = (2 pi)^(3/2)
    assert abs(L.RadialGrid_integrate(rad, 1, np.exp(-rp ** 2 / 2)) - (2 * np.pi) ** 1.5) < 1e-6
    d2 = np.ctypeslib.as_array(L.RadialGrid_deriv2(rad, np.exp(-rp ** 2)), shape=(50,)).copy()
    assert np.isfinite(d2).all()
    ang = L.AngularGrid_new(110)
    assert L.AngularGrid_get_lorder(ang) == 110
    u = np.ctypeslib.as_array(L.AngularGrid_get_points(ang), shape=(110, 3)).copy()
    aw = np.ctypeslib.as_array(L.AngularGrid_get_weights(ang), shape=(110,))
    assert abs(aw.sum() - 1.0) < 1e-14
    assert abs(L.AngularGrid_integrate(ang, 1, np.ones(110)) - 4 * np.pi) < 1e-13
    origin = np.array([0.3, -0.1, 0.2])
    at = L.AtomicGrid_new(ang, rad, origin)
    assert L.AtomicGrid_get_size(at) == 5500 and L.AtomicGrid_get_radii(at) == 0.5858151015155881
    np.testing.assert_array_equal(np.ctypeslib.as_array(L.AtomicGrid_get_origin(at), shape=(3,)), origin)
    ap = np.ctypeslib.as_array(L.AtomicGrid_get_points(at), shape=(5500, 3))
    awt = np.ctypeslib.as_array(L.AtomicGrid_get_weights(at), shape=(5500,))
    np.testing.assert_allclose(ap[110 * 3 + 5], origin + rp[3] * u[5], rtol=0, atol=1e-15)
    f = np.ascontiguousarray(np.exp(-((ap - origin) ** 2).sum(1)))
    seg = np.ctypeslib.as_array(L.AtomicGrid_integrate(at, 1, 1, 5500, f * awt), shape=(1,))
    assert abs(seg[0] - np.pi ** 1.5) < 1e-6
    # AngularGrid spherical expansion / evaluation on the unit sphere: the round trip of a
    # band-limited function is exact; Y_00 = 1 / sqrt(4 pi)
    lmax = 4
    L2 = (lmax + 1) ** 2
    fu = np.ascontiguousarray(1.0 + 0.5 * u[:, 2] + 0.25 * u[:, 0] * u[:, 1])
    ex = np.zeros((1, L2))
    L.AngularGrid_spherical_expansion(ang, lmax, 1, fu, ex)
    assert abs(ex[0, 0] - np.sqrt(4 * np.pi)) < 1e-12
    back = np.zeros(110)
    L.AngularGrid_eval_expansion(ang, lmax, 1, np.ascontiguousarray(ex), back)
    np.testing.assert_allclose(back, fu, rtol=0, atol=1e-12)
    L.AngularGrid_spherical(ang)
    sph = np.ctypeslib.as_array(L.AngularGrid_get_points(ang), shape=(110, 3))
    assert (sph[:, 0] == 1.0).all()
    np.testing.assert_allclose(np.cos(sph[:, 1]), u[:, 2], atol=1e-14)
    # FE assembly, cell, multigrid bookkeeping
    gold = np.load(os.path.join(systems.ROOT, 'tests', 'golden', 'ref_vectors.npz'))
    n = 12
    coeffs, rhs = np.zeros((2 * n, 2 * n)), np.zeros(2 * n)
    L.build_ode2(gold['ode_b'], gold['ode_a'], gold['ode_f'], gold['ode_bcs'], coeffs, rhs, n)
    np.testing.assert_array_equal(coeffs, gold['ode_coeffs'])
    cell = L.Cell_new(None, 0)
    L.Cell_del(cell)
    mg = L.MultiGrid_new(2)
    assert L.MultiGrid_get_ngrids(mg) == 0 and L.MultiGrid_get_nspecies(mg) == 2
    L.MultiGrid_del(mg)
    L.AtomicGrid_del(at)
    L.AngularGrid_del(ang)
    L.RadialGrid_del(rad)
    # every grid-path symbol that needs no GPU has been called above
    gpu_only = {'AtomicGrid_spherical_expansion', 'BeckeGrid_new', 'BeckeGrid_del', 'BeckeGrid_from_points',
                'BeckeGrid_integrate', 'BeckeGrid_get_ncenter', 'BeckeGrid_get_size', 'BeckeGrid_get_radii',
                'BeckeGrid_get_points', 'BeckeGrid_get_weights', 'BeckeGrid_get_origin',
                'BeckeGrid_get_becke_weights', 'eval_decomposition_grid', 'dot_multi_moments',
                'BasisSet_compute', 'BasisSet_deriv', 'MultiGrid_add_grid', 'MultiGrid_get_common_index',
                'MultiGrid_get_common_index_size', 'AuxiliaryBasis_new', 'AuxiliaryBasis_del',
                'AuxiliaryBasis_get_aobasis', 'AuxiliaryBasis_get_nao', 'nwavefunction_compute_2body_matrix',
                'nwavefunction_compute_coupling', 'nwavefunction_compute_xc_matrix',
                'nwavefunction_compute_density_from_dm', 'CubicSpline_del', 'RTransform_del'}
    missed = set(reftable.grid_path_symbols()) - L.calls - gpu_only
    assert missed == set(), sorted(missed)


def test_spline_derivative_is_the_derivative_of_the_spline():
    # independent of the golden vectors: central differences of CubicSpline_eval
    L = reftable.Bound()
    g = GOLD
    n = len(g['sp_r'])
    rtf = L.ChebyshevRTransform_new(0.8, n)
    y, dt = np.ascontiguousarray(g['sp_y']), np.zeros(n)
    L.solve_cubic_spline_system(y, dt, n)
    cs = L.CubicSpline_new(y, dt, L.PotentialExtrapolation_new(2), rtf, n)
    r = g['sp_r']
    x = np.ascontiguousarray(0.5 * (r[5:-5] + r[6:-4]))  # mid-interval points
    h = 1e-6 * x
    fp, fm, d = np.zeros(len(x)), np.zeros(len(x)), np.zeros(len(x))
    L.CubicSpline_eval(cs, np.ascontiguousarray(x + h), fp, len(x))
    L.CubicSpline_eval(cs, np.ascontiguousarray(x - h), fm, len(x))
    L.CubicSpline_eval_deriv(cs, x, d, len(x))
    np.testing.assert_allclose(d, (fp - fm) / (2 * h), rtol=1e-6, atol=1e-9)
