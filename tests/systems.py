"""Test and benchmark systems: the same geometry / basis / density on both sides
(napmo_amd on the GPU, oracle on the CPU). Seeded with numpy.random.default_rng."""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

from napmo_amd.workloads import (BASIS, H2_ATOMS, chon16_system, density_matrix, h2_system,  # noqa: F401,E402
                                named_shells, occupied_orbitals, random_molecule, random_shells)


def oracle_system(atoms, shells, nrad, nang, sizes=None):
    """(oracle BeckeGrid, oracle Basis) for the same system."""
    from oracle import core
    grid = core.BeckeGrid(atoms, nrad, nang, sizes=sizes)
    basis = core.Basis()
    for (s, o), sh in zip(atoms, shells):
        for shell in sh:
            basis.add_shell(o, shell['l'], shell['prim'], shell['cont'])
    return grid, basis


def gpu_system(atoms, shells, nrad, nang, sizes=None):
    import napmo_amd
    grid = napmo_amd.BeckeGrid(napmo_amd.make_species(atoms), nrad, nang, sizes=sizes)
    basis = napmo_amd.make_basis(atoms, shells_per_atom=shells)
    return grid, basis


def lebedev_lmax(nang):
    from oracle.core import lebedev_get_order
    return int(lebedev_get_order(nang) / 2)
