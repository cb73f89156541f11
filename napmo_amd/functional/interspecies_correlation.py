"""Interspecies correlation on the GPU (napmo/functional/interspecies_correlation.py:12-66)
and its application on the common points of two species' grids
(napmo/wavefunction/psi_analytic.py:214-262, compute_c_2species_grid)."""
import numpy as np

from napmo_amd.system import cext


def isc_functional_selector(name):
    if name is not None:
        assert isinstance(name, str)
        functional = {'epc17-2': epc17_2}.get(name)
        if functional is None:
            raise NotImplementedError(name + ' Functional NOT available!')
        return functional
    return None


def epc17_2(rhoE, rhoN):
    """epc17-2 (Yang, Brorsen, Culpitt, Pak, Hammes-Schiffer 2017): returns (c_ene, c_pot_A,
    c_pot_B), c_ene = -rhoN / (2.35 - 2.4 sqrt(rhoE rhoN) + 6.6 rhoE rhoN) and the
    derivatives of rhoE c_ene with respect to rhoE and rhoN."""
    rhoE = np.ascontiguousarray(rhoE, dtype=np.float64)
    rhoN = np.ascontiguousarray(rhoN, dtype=np.float64)
    assert rhoE.shape == rhoN.shape
    n = rhoE.size
    ene, pa, pb = np.zeros(n), np.zeros(n), np.zeros(n)
    cext.check(cext.napmo_epc17_2(n, rhoE.ctypes.data, rhoN.ctypes.data, ene.ctypes.data,
                                  pa.ctypes.data, pb.ctypes.data))
    return ene.reshape(rhoE.shape), pa.reshape(rhoE.shape), pb.reshape(rhoE.shape)


def c_2species_grid(grid_a, grid_b, index, rho_a, rho_b, vrho_a, vrho_b):
    """epc17-2 on the common points ``index`` [n x 2] (column 0 on grid_a, column 1 on
    grid_b) of the densities rho_a, rho_b given on the two grids: vrho_a, vrho_b (float64
    arrays on the grids) are accumulated in place; returns the correlation energy
    sum bw_b w_b c_ene rho_a, which the reference adds to both species."""
    index = np.ascontiguousarray(index, dtype=np.int32).reshape(-1, 2)
    rho_a = np.ascontiguousarray(rho_a, dtype=np.float64)
    rho_b = np.ascontiguousarray(rho_b, dtype=np.float64)
    assert rho_a.size == grid_a.size and rho_b.size == grid_b.size
    for v, g in ((vrho_a, grid_a), (vrho_b, grid_b)):
        assert v.dtype == np.float64 and v.flags.c_contiguous and v.size == g.size
    e = np.zeros(1)
    cext.check(cext.napmo_epc17_2_grid(grid_a._this, grid_b._this, index.ctypes.data, index.shape[0],
                                       rho_a.ctypes.data, rho_b.ctypes.data, vrho_a.ctypes.data,
                                       vrho_b.ctypes.data, e.ctypes.data))
    return float(e[0])
