"""napmo_amd: MI355X (gfx950) implementation of the nAPMO grid hot path.

Drop-in for the reference's grid Python API (napmo.grids, napmo.gto collocation,
the Poisson solver and the grid contractions) over libnapmo_gpu.so, plus the
device-resident grid Fock build ``GridFock``. Importing requires the built library
(``python napmo_amd/build.py``); there is no CPU fallback.
"""
from napmo_amd.system import cext  # noqa: F401  (raises if the library is missing)
from napmo_amd.grids import (AngularGrid, AtomicGrid, BeckeGrid, ChebyshevRadialTransform,  # noqa: F401
                             CubicSpline, CuspExtrapolation, IndentityRadialTransform,
                             PotentialExtrapolation, PowerExtrapolation, PowerRadialTransform,
                             RadialGrid, ZeroExtrapolation, lebedev_get_order, poisson_solver)
from napmo_amd.gto import BasisSet, ContractedGaussian, PrimitiveGaussian  # noqa: F401
from napmo_amd.utilities.ode2 import solve_ode2  # noqa: F401
from napmo_amd.utilities.cell import Cell  # noqa: F401
from napmo_amd.wavefunction.ntwobody import compute_coulomb  # noqa: F401
from napmo_amd.wavefunction.nkinetic import compute_kinetic  # noqa: F401
from napmo_amd.wavefunction.nnuclear import compute_nuclear  # noqa: F401
from napmo_amd.wavefunction.nexchange import compute_exchange_potential  # noqa: F401
from napmo_amd.wavefunction.ndpsi import compute_dpsi, becke_dickson_step  # noqa: F401
from napmo_amd.grids.multi_grid import MultiGrid  # noqa: F401
from napmo_amd.grids.auxiliary_basis import AuxiliaryBasis  # noqa: F401
from napmo_amd.functional import isc_functional_selector, epc17_2  # noqa: F401
from napmo_amd.wavefunction.psi_base import PSIB  # noqa: F401
from napmo_amd.system.elements import AtomicElementsDatabase  # noqa: F401
from napmo_amd.system.species import make_basis, make_species  # noqa: F401
from napmo_amd.fock import GridFock  # noqa: F401
from napmo_amd.apmo import APMOFock  # noqa: F401
from napmo_amd.batch import BatchFock  # noqa: F401
