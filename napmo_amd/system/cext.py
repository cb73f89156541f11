"""ctypes binding of libnapmo_gpu.so, the drop-in for the reference's ``libnapmo.so``
(napmo/system/cext.py:11-16; argtypes table of napmo/__init__.py:95-585 for the grid
hot path, plus the device-resident ``napmo_*`` entry points of include/napmo_gpu.h).

There is no CPU fallback: importing this module without the built library raises.
"""
import ctypes
import os

import numpy as np
import numpy.ctypeslib as npct

_HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(_HERE, '..', 'libnapmo_gpu.so')

if not os.path.exists(LIBPATH):
    raise ImportError('napmo_amd: %s is not built (run `python napmo_amd/build.py`); '
                      'there is no CPU fallback' % os.path.normpath(LIBPATH))

lib = ctypes.CDLL(LIBPATH)

# array types (napmo/__init__.py:96-102)
a1df = npct.ndpointer(dtype=np.double, ndim=1, flags='C_CONTIGUOUS')
a2df = npct.ndpointer(dtype=np.double, ndim=2, flags='C_CONTIGUOUS')
a3df = npct.ndpointer(dtype=np.double, ndim=3, flags='C_CONTIGUOUS')
adf = npct.ndpointer(dtype=np.double, flags='C_CONTIGUOUS')
a1di = npct.ndpointer(dtype=np.int32, ndim=1, flags='C_CONTIGUOUS')
a1dli = npct.ndpointer(dtype=np.int64, ndim=1, flags='C_CONTIGUOUS')
aptr = npct.ndpointer(dtype=ctypes.c_void_p, flags='C_CONTIGUOUS')
vp = ctypes.c_void_p
c_int, c_long, c_double = ctypes.c_int, ctypes.c_long, ctypes.c_double
dptr = ctypes.POINTER(ctypes.c_double)


class WaveFunction(ctypes.Structure):
    """src/wavefunction/wavefunction.h:20-44 (== napmo/wavefunction/psi_base.py:22-45)."""
    _fields_ = [(name, dptr) for name in ('_S', '_X', '_T', '_V', '_H', '_C', '_D', '_L', '_G', '_J',
                                          '_XC', '_F', '_O')] + [
        ('_nbasis', c_int), ('_ndim', c_int), ('_occupation', c_int), ('_eta', c_double),
        ('_kappa', c_double), ('_x_factor', c_double), ('_xc_energy', c_double),
        ('_energy', c_double), ('_rmsd', c_double)]


_SIGNATURES = {
    # errors / memory
    'napmo_last_error': (ctypes.c_char_p, []),
    'napmo_free': (None, [vp]),
    'napmo_device_count': (c_int, []),
    # radial transforms
    'ChebyshevRTransform_new': (vp, [c_double, c_int]),
    'ChebyshevRTransform_get_radii': (c_double, [vp]),
    'ChebyshevRTransform_get_z': (c_double, [vp]),
    'PowerRTransform_new': (vp, [c_double, c_double, c_int]),
    'PowerRTransform_get_rmin': (c_double, [vp]),
    'PowerRTransform_get_rmax': (c_double, [vp]),
    'PowerRTransform_get_power': (c_double, [vp]),
    'IdentityRTransform_new': (vp, [c_int]),
    'RTransform_get_npoint': (c_int, [vp]),
    'RTransform_radius': (c_double, [vp, c_double]),
    'RTransform_deriv': (c_double, [vp, c_double]),
    'RTransform_deriv2': (c_double, [vp, c_double]),
    'RTransform_deriv3': (c_double, [vp, c_double]),
    'RTransform_inv': (c_double, [vp, c_double]),
    'RTransform_radius_array': (None, [vp, a1df, a1df, c_int]),
    'RTransform_deriv_array': (None, [vp, a1df, a1df, c_int]),
    'RTransform_deriv2_array': (None, [vp, a1df, a1df, c_int]),
    'RTransform_deriv3_array': (None, [vp, a1df, a1df, c_int]),
    'RTransform_inv_array': (None, [vp, a1df, a1df, c_int]),
    'RTransform_del': (None, [vp]),
    # grids
    'AngularGrid_new': (vp, [c_int]),
    'AngularGrid_get_lorder': (c_int, [vp]),
    'AngularGrid_get_points': (dptr, [vp]),
    'AngularGrid_get_weights': (dptr, [vp]),
    'AngularGrid_integrate': (c_double, [vp, c_int, a1df]),
    'AngularGrid_del': (None, [vp]),
    'AngularGrid_spherical': (None, [vp]),
    'AngularGrid_spherical_expansion': (None, [vp, c_int, c_int, a1df, a2df]),
    'AngularGrid_eval_expansion': (None, [vp, c_int, c_int, a2df, a1df]),
    'RadialGrid_new': (vp, [vp, c_double]),
    'RadialGrid_get_size': (c_int, [vp]),
    'RadialGrid_get_radii': (c_double, [vp]),
    'RadialGrid_get_points': (dptr, [vp]),
    'RadialGrid_get_weights': (dptr, [vp]),
    'RadialGrid_integrate': (c_double, [vp, c_int, a1df]),
    'RadialGrid_deriv2': (dptr, [vp, a1df]),
    'RadialGrid_del': (None, [vp]),
    'AtomicGrid_new': (vp, [vp, vp, a1df]),
    'AtomicGrid_get_size': (c_int, [vp]),
    'AtomicGrid_get_radii': (c_double, [vp]),
    'AtomicGrid_get_origin': (dptr, [vp]),
    'AtomicGrid_get_points': (dptr, [vp]),
    'AtomicGrid_get_weights': (dptr, [vp]),
    'AtomicGrid_spherical_expansion': (dptr, [vp, a1df, c_int]),
    'AtomicGrid_integrate': (dptr, [vp, c_int, c_int, c_int, a1df]),
    'AtomicGrid_del': (None, [vp]),
    'BeckeGrid_new': (vp, [aptr, c_int, c_int, c_double]),
    'BeckeGrid_get_ncenter': (c_int, [vp]),
    'BeckeGrid_get_size': (c_int, [vp]),
    'BeckeGrid_get_radii': (dptr, [vp]),
    'BeckeGrid_get_points': (dptr, [vp]),
    'BeckeGrid_get_weights': (dptr, [vp]),
    'BeckeGrid_get_origin': (dptr, [vp]),
    'BeckeGrid_get_becke_weights': (dptr, [vp]),
    'BeckeGrid_integrate': (c_double, [vp, a1df]),
    'BeckeGrid_del': (None, [vp]),
    'BeckeGrid_from_points': (vp, [a2df, c_int, c_int]),
    'AuxiliaryBasis_new': (vp, [vp]),
    'AuxiliaryBasis_del': (None, [vp]),
    'AuxiliaryBasis_compute_aobasis': (None, [vp]),
    'AuxiliaryBasis_get_aobasis': (dptr, [vp]),
    'AuxiliaryBasis_get_nao': (c_int, [vp]),
    'AuxiliaryBasis_get_ncenter': (c_int, [vp]),
    'MultiGrid_new': (vp, [c_int]),
    'MultiGrid_del': (None, [vp]),
    'MultiGrid_add_grid': (c_int, [vp, vp]),
    'MultiGrid_get_common_index': (ctypes.POINTER(c_int), [vp, c_int, c_int]),
    'MultiGrid_get_common_index_size': (c_int, [vp, c_int, c_int]),
    'MultiGrid_get_ngrids': (c_int, [vp]),
    'MultiGrid_get_nspecies': (c_int, [vp]),
    # splines, extrapolation, ODE
    'solve_cubic_spline_system': (None, [a1df, a1df, c_int]),
    'CubicSpline_new': (vp, [a1df, a1df, vp, vp, c_int]),
    'CubicSpline_eval': (None, [vp, a1df, a1df, c_int]),
    'CubicSpline_eval_deriv': (None, [vp, a1df, a1df, c_int]),
    'CubicSpline_get_rtransform': (vp, [vp]),
    'CubicSpline_get_extrapolation': (vp, [vp]),
    'compute_cubic_spline_int_weights': (None, [a1df, c_int]),
    'CubicSpline_get_first_x': (c_double, [vp]),
    'CubicSpline_get_last_x': (c_double, [vp]),
    'CubicSpline_del': (None, [vp]),
    'PotentialExtrapolation_new': (vp, [c_int]),
    'ZeroExtrapolation_new': (vp, []),
    'CuspExtrapolation_new': (vp, []),
    'PowerExtrapolation_new': (vp, [c_double]),
    'Extrapolation_del': (None, [vp]),
    'Extrapolation_prepare': (None, [vp, vp]),
    'Extrapolation_eval_left': (c_double, [vp, c_double]),
    'Extrapolation_eval_right': (c_double, [vp, c_double]),
    'Extrapolation_deriv_left': (c_double, [vp, c_double]),
    'Extrapolation_deriv_right': (c_double, [vp, c_double]),
    'Extrapolation_has_tail': (ctypes.c_bool, [vp]),
    'PowerExtrapolation_get_power': (c_double, [vp]),
    'build_ode2': (None, [a1df, a1df, a1df, a1df, a2df, a1df, c_long]),
    'Cell_new': (vp, [vp, c_int]),
    'Cell_del': (None, [vp]),
    'eval_decomposition_grid': (None, [aptr, a1df, a1df, a2df, vp, c_long, c_long]),
    'dot_multi_moments': (None, [c_long, c_long, a1df, a2df, a1df, c_long, c_long, a1dli, a2df, c_long]),
    # Gaussian basis
    'PrimitiveGaussian_new': (vp, [a1di, a1df, c_double, c_double]),
    'PrimitiveGaussian_get_norma': (c_double, [vp]),
    'PrimitiveGaussian_compute': (c_double, [vp, a2df, a1df, c_int]),
    'PrimitiveGaussian_overlap': (c_double, [vp, vp]),
    'PrimitiveGaussian_get_l': (None, [vp, a1di]),
    'PrimitiveGaussian_get_origin': (None, [vp, a1df]),
    'PrimitiveGaussian_get_zeta': (c_double, [vp]),
    'PrimitiveGaussian_get_coeff': (c_double, [vp]),
    'ContractedGaussian_new': (vp, [aptr, c_int]),
    'ContractedGaussian_get_norma': (c_double, [vp]),
    'ContractedGaussian_get_nprim': (c_int, [vp]),
    'ContractedGaussian_get_l': (None, [vp, a1di]),
    'ContractedGaussian_get_origin': (None, [vp, a1df]),
    'ContractedGaussian_compute': (None, [vp, a2df, a1df, c_int]),
    'ContractedGaussian_overlap': (c_double, [vp, vp]),
    'BasisSet_new': (vp, [aptr, c_int]),
    'BasisSet_new_empty': (vp, []),
    'BasisSet_update': (None, [vp, vp]),
    'BasisSet_get_nbasis': (c_int, [vp]),
    'BasisSet_compute': (None, [vp, a2df, a2df, c_int]),
    'BasisSet_deriv': (None, [vp, a2df, a3df, c_int]),
    'BasisSet_get_max_l': (c_int, [vp]),
    'BasisSet_get_max_nprim': (c_int, [vp]),
    # contractions
    'nwavefunction_compute_2body_matrix': (None, [ctypes.POINTER(WaveFunction), vp, a2df, a1df, a2df]),
    'nwavefunction_compute_coupling': (None, [ctypes.POINTER(WaveFunction), vp, a2df, a1df, a2df]),
    'nwavefunction_compute_xc_matrix': (None, [ctypes.POINTER(WaveFunction), vp, a2df, a1df]),
    'nwavefunction_compute_density_from_dm': (None, [vp, vp, a2df, a1df, c_double, a1df]),
    # device-resident hot path
    'napmo_compute_coulomb': (c_int, [vp, a1df, c_int, a1df]),
    'napmo_poisson_solve': (c_int, [vp, a1df, c_int, adf, adf]),
    'napmo_poisson_solve_expansion': (c_int, [vp, adf, c_int, adf, adf]),
    'napmo_compute_kinetic': (c_int, [vp, a1df, c_int, a1df]),
    'napmo_kinetic_expand': (c_int, [vp, a1df, c_int, adf, adf, vp]),
    'napmo_compute_nuclear': (c_int, [vp, c_int, a1df, a2df, a1df]),
    'napmo_compute_exchange_potential': (c_int, [vp, c_int, a2df, a2df, c_int, a2df]),
    'napmo_compute_exchange_potential_device': (c_int, [vp, c_int, vp, vp, c_int, vp]),
    'napmo_compute_kinetic_device': (c_int, [vp, vp, c_int, vp]),
    'napmo_compute_dpsi': (c_int, [vp, c_int, c_int, vp, vp, vp, vp, vp, c_double, vp]),
    'napmo_becke_dickson_step': (c_int, [vp, c_int, c_int, vp, vp, vp, vp, c_double, vp, vp, vp]),
    'napmo_becke_dickson_step_device': (c_int, [vp, c_int, c_int, vp, vp, vp, vp, c_double, vp,
                                                vp, vp]),
    'napmo_grid_profile': (c_int, [vp, c_int]),
    'napmo_epc17_2': (c_int, [c_long, vp, vp, vp, vp, vp]),
    'napmo_epc17_2_grid': (c_int, [vp, vp, vp, c_int, vp, vp, vp, vp, vp]),
    'napmo_epc17_2_grid_noalloc': (c_int, [vp, vp, vp, c_int, vp, vp, vp, vp, vp]),
    'napmo_grid_sync': (c_int, [vp]),
    'napmo_grid_profile_read': (c_int, [vp, vp, c_int, ctypes.POINTER(c_int)]),
    'napmo_fock_plan_new': (c_int, [vp, vp, c_int, c_int, ctypes.POINTER(vp)]),
    'napmo_fock_plan_del': (None, [vp]),
    'napmo_fock_plan_info': (c_int, [vp, ctypes.POINTER(c_long), ctypes.POINTER(c_int),
                                     ctypes.POINTER(c_int), ctypes.POINTER(c_int)]),
    'napmo_grid_fock_build': (c_int, [vp, vp, vp, vp, vp, vp]),
    'napmo_fock_plan_centres': (c_int, [vp, ctypes.POINTER(c_int), vp]),
    'napmo_grid_fock_gradient': (c_int, [vp, vp, vp]),
    'napmo_grid_fock_gradient_device': (c_int, [vp, vp, vp]),
    'napmo_grid_fock_energy_dP': (c_int, [vp, vp]),
    'napmo_grid_fock_energy_dP_device': (c_int, [vp, vp]),
    'napmo_grid_poisson_adjoint': (c_int, [vp, vp, vp]),
    'napmo_fock_plan_set_tuning': (c_int, [vp, c_int, c_long]),
    'napmo_fock_plan_get_tuning': (c_int, [vp, c_int, ctypes.POINTER(c_long)]),
    'napmo_fock_plan_int8_screen_stats': (c_int, [vp, ctypes.POINTER(c_long), ctypes.POINTER(c_long)]),
    'napmo_grid_set_tuning': (c_int, [vp, c_int, c_long]),
    'napmo_grid_get_tuning': (c_int, [vp, c_int, ctypes.POINTER(c_long)]),
    'napmo_batch_plan_new': (c_int, [c_int, vp, c_int, ctypes.POINTER(vp)]),
    'napmo_batch_plan_del': (None, [vp]),
    'napmo_batch_build': (c_int, [vp, vp, vp, vp]),
    'napmo_batch_capture': (c_int, [vp, vp, vp, vp]),
    'napmo_batch_build_orbitals': (c_int, [vp, vp, vp, c_double, vp, vp]),
    'napmo_batch_capture_orbitals': (c_int, [vp, vp, vp, c_double, vp, vp]),
    'napmo_batch_replay': (c_int, [vp]),
    'napmo_batch_sync': (c_int, [vp]),
    'napmo_apmo_plan_new': (c_int, [c_int, vp, vp, vp, vp, c_int, ctypes.POINTER(vp)]),
    'napmo_apmo_plan_del': (None, [vp]),
    'napmo_apmo_build': (c_int, [vp, vp, vp, vp]),
    'napmo_apmo_build_device': (c_int, [vp, vp, vp, vp]),
    'napmo_apmo_sync': (c_int, [vp]),
    'napmo_apmo_plan_set_one_launch': (c_int, [vp, c_int]),
    'napmo_apmo_plan_set_tuning': (c_int, [vp, c_int, c_int, c_long]),
    'napmo_apmo_capture': (c_int, [vp, vp, vp, vp]),
    'napmo_apmo_replay': (c_int, [vp]),
    'napmo_grid_fock_build_device': (c_int, [vp, vp, vp, vp]),
    'napmo_grid_fock_build_orbitals': (c_int, [vp, vp, c_int, c_double, vp, vp, vp, vp]),
    'napmo_grid_fock_build_orbitals_device': (c_int, [vp, vp, c_int, c_double, vp, vp]),
    'napmo_selftest_mfma': (c_int, [c_int, a1df]),
    'napmo_device_copy': (c_int, [vp, vp, c_long]),
    'napmo_grid_overlap': (c_int, [vp, a2df]),
    'napmo_fock_plan_contract': (c_int, [vp, a1df, a2df]),
    'napmo_fock_plan_profile': (c_int, [vp, c_int]),
    'napmo_fock_plan_profile_read': (c_int, [vp, a1df, c_int, ctypes.POINTER(c_int)]),
    'napmo_fock_phase_name': (ctypes.c_char_p, [c_int]),
    'napmo_device_malloc': (vp, [ctypes.c_size_t]),
    'napmo_device_free': (None, [vp]),
    'napmo_memcpy': (c_int, [vp, vp, ctypes.c_size_t, c_int]),
    'napmo_fock_plan_stream': (vp, [vp]),
    'napmo_fock_plan_sync': (c_int, [vp]),
}

for _name, (_res, _args) in _SIGNATURES.items():
    _fn = getattr(lib, _name)
    _fn.restype = _res
    _fn.argtypes = _args
    globals()[_name] = _fn

EXPORTED = tuple(_SIGNATURES)


class NapmoError(RuntimeError):
    pass


def check(status):
    """Raise NapmoError with the library's message for a non-zero status."""
    if status != 0:
        raise NapmoError(lib.napmo_last_error().decode())
    return status


def require_gpu():
    if lib.napmo_device_count() < 1:
        raise NapmoError('napmo_amd: no HIP device visible; the grid hot path runs only on the GPU')


def as_array(ptr, shape):
    """Borrowed view of a library-owned buffer (napmo/grids/becke.py:120-145)."""
    return np.ctypeslib.as_array(ptr, shape=shape)


def take_array(ptr, shape):
    """Copy of a callee-allocated buffer, then released with napmo_free."""
    out = np.ctypeslib.as_array(ptr, shape=shape).copy()
    lib.napmo_free(ctypes.cast(ptr, vp))
    return out
