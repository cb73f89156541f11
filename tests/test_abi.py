"""The C-ABI library loads, exports every symbol include/napmo_gpu.h declares, and its
host-side setup objects (transforms, Lebedev rules, radial/atomic grids, splines,
FE assembly) agree with the reference's routines. No GPU calls."""
import os
import re

import numpy as np
import pytest

from tests import systems

ROOT = systems.ROOT
GOLD = np.load(os.path.join(ROOT, 'tests', 'golden', 'ref_vectors.npz'))


def declared_symbols():
    text = open(os.path.join(ROOT, 'include', 'napmo_gpu.h')).read()
    text = re.sub(r'/\*.*?\*/', '', text, flags=re.S)
    text = re.sub(r'typedef struct.*?\}\s*\w+\s*;', '', text, flags=re.S)
    return sorted(set(re.findall(r'\b(\w+)\s*\([^;{]*\)\s*;', text)))


def test_library_exports_every_declared_symbol():
    from napmo_amd.system import cext
    names = declared_symbols()
    assert len(names) > 80
    missing = [n for n in names if not hasattr(cext.lib, n)]
    assert not missing, missing
    # and the Python binding table covers them all
    assert set(names) <= set(cext.EXPORTED), set(names) - set(cext.EXPORTED)


def test_lebedev_rules_match_reference():
    import napmo_amd
    for n in [6, 14, 26, 38, 50, 74, 86, 110, 146, 170, 194, 230, 266, 302, 350, 434, 590]:
        g = napmo_amd.AngularGrid(n)
        assert np.array_equal(g.points, GOLD['leb_%d_xyz' % n])
        assert np.array_equal(g.weights, GOLD['leb_%d_w' % n])
    with pytest.raises(KeyError):
        napmo_amd.AngularGrid(7)


def test_radial_transforms_match_reference():
    import napmo_amd
    t = np.arange(100, dtype=float)
    for name, rtf in (('cheb', napmo_amd.ChebyshevRadialTransform(0.5858151015155881, 100)),
                      ('power', napmo_amd.PowerRadialTransform(1e-3, 2e1, 100))):
        for which in ('radius', 'deriv', 'deriv2', 'deriv3'):
            ref = GOLD['rtf_%s_%s' % (name, which)]
            assert np.max(np.abs(getattr(rtf, which)(t) - ref) / np.abs(ref)) < 1e-14
        assert np.max(np.abs(rtf.inv(GOLD['rtf_%s_radius' % name]) - GOLD['rtf_%s_inv' % name])) < 1e-12
    assert napmo_amd.lebedev_get_order(110) == 17


def test_radial_and_atomic_grids_match_oracle():
    import napmo_amd
    from oracle import core
    rg = napmo_amd.RadialGrid(50, 'He')
    r, rw = core.radial_grid(core.ChebyshevRTransform(core.atomic_radius('He'), 50))
    assert np.array_equal(rg.points, r)
    assert np.allclose(rg.weights, rw, rtol=1e-15)
    ag = napmo_amd.AtomicGrid(20, 50, np.array([0.1, 0.2, -0.3]), 'O')
    og = core.AtomicGrid(20, 50, [0.1, 0.2, -0.3], 'O')
    assert np.array_equal(ag.points, og.points)
    assert np.allclose(ag.weights, og.weights, rtol=1e-15)
    # radial integral closed form (napmo/grids/test/test_radial.py)
    y = np.exp(-0.5 * rg.points ** 2)
    rg2 = napmo_amd.RadialGrid(100)
    assert abs(rg2.integrate(np.exp(-0.5 * rg2.points ** 2), 1) - (2 * np.pi) ** 1.5) < 1e-9
    del y


def test_splines_and_ode2_match_reference():
    import napmo_amd
    from napmo_amd.system import cext
    dt = np.zeros_like(GOLD['spline_y'])
    cext.solve_cubic_spline_system(np.ascontiguousarray(GOLD['spline_y']), dt, len(dt))
    assert np.array_equal(dt, GOLD['spline_dt'])
    coeffs = np.zeros((24, 24))
    rhs = np.zeros(24)
    cext.build_ode2(GOLD['ode_b'], GOLD['ode_a'], GOLD['ode_f'], GOLD['ode_bcs'], coeffs, rhs, 12)
    assert np.array_equal(coeffs, GOLD['ode_coeffs']) and np.array_equal(rhs, GOLD['ode_rhs'])
    # extrapolation behaviour (napmo/grids/test/test_extrapolation.py:14-36)
    x = np.arange(10, dtype=float)
    y = np.exp(-0.3 * x)
    cs = napmo_amd.CubicSpline(y, -0.3 * y)
    assert abs(cs(np.array([-2.5, -1.1])) - np.exp(-0.3 * np.array([-2.5, -1.1]))).max() < 1e-10
    assert abs(cs(np.array([10.5, 11.5]))).max() < 1e-10
    cs = napmo_amd.CubicSpline(1 / (np.exp(-2 * x) + x ** 2),
                               -(1 / (np.exp(-2 * x) + x ** 2)) ** 2 * (np.exp(-2 * x) * (-2) + 2 * x),
                               extrapolation=napmo_amd.PowerExtrapolation(-2.0))
    assert abs(cs(np.array([-2.5, -1.1]))).max() == 0.0
    assert abs(cs(np.array([10.5, 11.5])) - 1 / np.array([10.5, 11.5]) ** 2).max() < 1e-10


def test_solve_ode2_python_api_matches_oracle():
    import napmo_amd
    from oracle import core
    rtf = napmo_amd.ChebyshevRadialTransform(1.0, 30)
    r = rtf.radius_all()
    b = napmo_amd.CubicSpline(2 / r, -2 / r ** 2, rtransform=rtf)
    a = napmo_amd.CubicSpline(-2 / r ** 2, 4 / r ** 3, rtransform=rtf)
    rho = np.exp(-r)
    f = napmo_amd.CubicSpline(-4 * np.pi * rho, rtransform=rtf)
    out = napmo_amd.solve_ode2(b, a, f, (0.1, None, 0.0, None), napmo_amd.PotentialExtrapolation(1))
    ortf = core.ChebyshevRTransform(1.0, 30)
    t = np.arange(30.0)
    d1 = ortf.deriv(t)
    fdt = core.spline_dt(-4 * np.pi * rho)
    ref = core.solve_ode2((2 / r, -2 / r ** 2), (-2 / r ** 2, 4 / r ** 3), (-4 * np.pi * rho, fdt / d1),
                          (0.1, None, 0.0, None), ortf, d1, ortf.deriv2(t), ortf.deriv3(t),
                          core.PotentialExtrapolation(1))
    assert np.allclose(out.y, ref.y, rtol=1e-10, atol=1e-12)


def test_gto_normalisation_constants():
    import napmo_amd
    # napmo/gto/test/test_primitive_gaussian.py, test_contracted_gaussian.py
    p = napmo_amd.PrimitiveGaussian(0.5)
    assert abs(p.normalization - 0.423777208124) < 1e-12
    p3 = napmo_amd.PrimitiveGaussian(0.5, l=np.array([0, 0, 3], dtype=np.int32))
    assert abs(p3.normalization - 0.309483114995) < 1e-12
    c = napmo_amd.ContractedGaussian(exponents=[0.5], l=np.array([0, 0, 3], dtype=np.int32))
    assert abs(c.compute(np.ones([1, 3]))[0] - 0.069055017012) < 1e-12
    with pytest.raises(ValueError):
        napmo_amd.BasisSet('STO-3G', 'HH', origin=[0.0, 0.0, 0.0])


def test_radial_deriv2_matches_oracle():
    # RadialGrid_deriv2 is host code (legacy export): same restatement as the oracle
    import napmo_amd
    from oracle import core
    rt = napmo_amd.ChebyshevRadialTransform(0.8, 40)
    rg = napmo_amd.RadialGrid(40, 'H', rt)
    r = np.array(rg.points)
    f = np.exp(-0.9 * r * r) + 0.1 * r * np.exp(-0.2 * r)
    mine = rg.deriv2(f)
    ref = core.radial_deriv2(rg.radii, r, f)
    assert np.max(np.abs(mine - ref)) <= 1e-13 * np.abs(ref).max()
