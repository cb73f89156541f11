"""Writes tests/golden/ctypes_table.json: the reference's ctypes declarations
(napmo/__init__.py:95-585), read as text, one entry per symbol: its restype and argtypes
as the source spells them, and whether the symbol belongs to the grid hot path that
libnapmo_gpu.so replaces (everything except the Libint interface and the host SCF
driver's wavefunction_* calls, which stay in the reference's own library).

Usage: python tests/golden/make_ctypes_table.py [/root/reference]
"""
import json
import os
import re
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
OUT_OF_PATH = ('LibintInterface_', 'wavefunction_')


def main(ref_root):
    text = open(os.path.join(ref_root, 'napmo', '__init__.py')).read()
    text = re.sub(r'#[^\n]*', '', text)  # commented-out declarations are not bound
    table = {}
    for m in re.finditer(r'cext\.(\w+)\.(restype|argtypes)\s*=\s*(\[[^\]]*\]|[^\n]+)', text):
        name, what, val = m.group(1), m.group(2), ' '.join(m.group(3).split())
        ent = table.setdefault(name, {'restype': 'int', 'argtypes': None})
        if what == 'restype':
            ent['restype'] = val
        else:
            ent['argtypes'] = [a.strip() for a in val.strip('[]').split(',') if a.strip()]
    for name, ent in table.items():
        ent['grid_path'] = not name.startswith(OUT_OF_PATH)
    with open(os.path.join(HERE, 'ctypes_table.json'), 'w') as f:
        json.dump({'source': 'napmo/__init__.py:95-585 (ctypes declarations, read as text)',
                   'symbols': dict(sorted(table.items()))}, f, indent=1)
    print(len(table), 'symbols,', sum(e['grid_path'] for e in table.values()), 'on the grid path')


if __name__ == '__main__':
    main(sys.argv[1] if len(sys.argv) > 1 else '/root/reference')
