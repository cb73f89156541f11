"""Regenerates ``elements.json`` and ``basis_sets.json`` from the reference's data files.

Both are published tabulations, kept here as plain numbers:

* atomic radii (bohr), atomic numbers and masses, as the reference's
  ``napmo/data/atomic_elements.json`` holds them (radii are what
  ``napmo/grids/radial.py:20-48`` uses for the Chebyshev radial transform);
* contracted Gaussian basis sets from the EMSL Basis Set Exchange, as the
  reference's deMon2k-format files under ``napmo/data/basis/`` hold them
  (format read by ``napmo/gto/basis_set.py:56-120``).

Usage: ``python napmo_amd/data/make_data_tables.py [/root/reference]``.
"""
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, '..', '..'))

from napmo_amd.gto.basis_io import parse_demon2k  # noqa: E402

BASIS_SETS = ['STO-3G', 'CC-PVDZ', '6-31G']
ELEMENTS = ['H', 'He', 'Li', 'Be', 'B', 'C', 'N', 'O', 'F', 'Ne']


def main(ref_root):
    with open(os.path.join(ref_root, 'napmo', 'data', 'atomic_elements.json')) as f:
        db = json.load(f)
    elements = {}
    for sym, rec in db.items():
        if not isinstance(rec, dict) or 'atomic_radii' not in rec:
            continue
        elements[sym] = {k: rec.get(k) for k in
                         ('atomic_number', 'atomic_radii', 'atomic_radii_2', 'mass', 'charge')}
    with open(os.path.join(HERE, 'elements.json'), 'w') as f:
        json.dump(elements, f, indent=0, sort_keys=True)

    basis = {}
    for name in BASIS_SETS:
        path = os.path.join(ref_root, 'napmo', 'data', 'basis', name)
        with open(path) as f:
            text = f.read()
        basis[name] = {}
        for sym in ELEMENTS:
            try:
                basis[name][sym] = parse_demon2k(text, sym)
            except ValueError:
                pass
    with open(os.path.join(HERE, 'basis_sets.json'), 'w') as f:
        json.dump(basis, f, indent=0, sort_keys=True)


if __name__ == '__main__':
    main(sys.argv[1] if len(sys.argv) > 1 else '/root/reference')
