"""Species records for the grid API (the part of napmo/system/molecular_system.py that
BeckeGrid and BasisSet read): {'symbol', 'particles': [{'symbol', 'origin'}]}."""
import numpy as np


def make_species(atoms, symbol='e-'):
    """atoms: [(element_symbol, (x, y, z) in bohr), ...]."""
    return {'symbol': symbol,
            'particles': [{'symbol': s, 'origin': np.asarray(o, dtype=np.float64)} for s, o in atoms]}


def make_basis(atoms, basis_name=None, shells_per_atom=None):
    """One BasisSet over all centres: from the packaged tables (basis_name) or explicit
    shells [[{'l', 'prim', 'cont'}, ...] per atom]."""
    from napmo_amd.gto.basis_set import BasisSet
    total = None
    for k, (s, o) in enumerate(atoms):
        shells = None if shells_per_atom is None else shells_per_atom[k]
        if shells is not None and len(shells) == 0:
            continue
        b = BasisSet(basis_name or 'custom', s, origin=np.asarray(o, dtype=np.float64), shells=shells)
        if total is None:
            total = b
        else:
            total.update(b)
    return total
