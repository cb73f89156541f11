"""Regenerates ``lebedev_orbits.json``: the Lebedev-Laikov octahedral quadrature rules.

The rules are published constants (V. Lebedev, D. Laikov, Doklady Mathematics 59(3),
477-481, 1999). The reference ships them as generator code
(src/grid/sphere_lebedev_rule.cpp:61 ``gen_oh``, :639 ``ld_by_order``). This script
runs that code once, from oracle/_ref/libnapmo_ref.so (built by ``make -C oracle ref``
from the reference sources where they lie), and reduces its output to this
package's own compact form:

* ``templates[code]``: the ordered list of signed coordinate patterns that one
  octahedral orbit of type ``code`` (1..6) expands to. Point order inside a rule
  follows the reference, so that atomic-grid point indices agree.
* ``rules[npoint]``: the orbits of each rule as ``[code, a, b, v]``.

Only the numbers are kept; ``napmo_amd/csrc`` and ``oracle/`` expand them with their
own code. Usage: ``python napmo_amd/data/make_lebedev_table.py`` (needs oracle/_ref).
"""
import ctypes
import json
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORDERS = [6, 14, 26, 38, 50, 74, 86, 110, 146, 170, 194, 230, 266, 302, 350, 434,
          590, 770, 974, 1202, 1454, 1730, 2030, 2354, 2702, 3074, 3470, 3890, 4334,
          4802, 5294, 5810]
ORBIT_SIZE = {1: 6, 2: 12, 3: 8, 4: 24, 5: 24, 6: 48}


def _ref_rule(lib, n):
    out = [np.zeros(n) for _ in range(4)]
    lib.ref_ld_by_order(n, *out)
    return np.stack(out[:3], axis=1), out[3]


def _classify(pts):
    """Orbit code from the multiset of |coordinates| of one orbit's first point."""
    p = np.abs(pts[0])
    nz = int((p == 0.0).sum())
    if nz == 2:
        return 1
    if nz == 1:
        vals = p[p != 0.0]
        return 2 if vals[0] == vals[1] else 5
    if p[0] == p[1] == p[2]:
        return 3
    if p[0] == p[1] or p[0] == p[2] or p[1] == p[2]:
        return 4
    return 6


def _params(code, first):
    a = b = 0.0
    f = np.abs(first)
    if code == 4:  # (a, a, b): a is the repeated value
        a = f[0] if (f[0] == f[1] or f[0] == f[2]) else f[1]
    elif code == 5:  # first point is (a, b, 0)
        a = f[0]
    elif code == 6:  # first point is (a, b, c)
        a, b = f[0], f[1]
    return float(a), float(b)


def _symbols(code, a, b):
    """Magnitudes of the symbols A, B, C of one orbit, computed as the expansion does."""
    if code == 1:
        return 1.0, 0.0, 0.0
    if code == 2:
        return 1.0 / np.sqrt(2.0), 0.0, 0.0
    if code == 3:
        return 1.0 / np.sqrt(3.0), 0.0, 0.0
    if code == 4:
        return a, float(np.sqrt(1.0 - 2.0 * a * a)), 0.0
    if code == 5:
        return a, float(np.sqrt(1.0 - a * a)), 0.0
    return a, b, float(np.sqrt(1.0 - a * a - b * b))


def _template(code, pts, a, b):
    sym = _symbols(code, a, b)
    tmpl = []
    for p in pts:
        row = []
        for v in p:
            if v == 0.0:
                row.append(0)
                continue
            k = [i for i in range(3) if sym[i] == abs(v)]
            assert len(k) == 1, (code, p, sym)
            row.append((k[0] + 1) * (1 if v > 0 else -1))
        tmpl.append(row)
    return tmpl


def main():
    lib = ctypes.CDLL(os.path.join(HERE, '..', '..', 'oracle', '_ref', 'libnapmo_ref.so'))
    ptr = np.ctypeslib.ndpointer(np.float64, flags='C')
    lib.ref_ld_by_order.argtypes = [ctypes.c_int, ptr, ptr, ptr, ptr]
    templates, rules = {}, {}
    for n in ORDERS:
        xyz, w = _ref_rule(lib, n)
        i, orbits = 0, []
        while i < n:
            code = _classify(xyz[i:i + 1])
            size = ORBIT_SIZE[code]
            chunk = xyz[i:i + size]
            assert np.all(w[i:i + size] == w[i])
            a, b = _params(code, chunk[0])
            t = _template(code, chunk, a, b)
            if code in templates:
                assert templates[code] == t, (n, code)
            else:
                templates[code] = t
            orbits.append([code, a, b, float(w[i])])
            i += size
        rules[str(n)] = orbits
    data = {
        'source': 'Lebedev & Laikov (1999) octahedral rules, point order as the reference '
                  'generator (src/grid/sphere_lebedev_rule.cpp)',
        'symbols': 'template entries: 0 = zero, +-1 = A, +-2 = B, +-3 = C',
        'templates': {str(k): v for k, v in sorted(templates.items())},
        'rules': rules,
    }
    with open(os.path.join(HERE, 'lebedev_orbits.json'), 'w') as f:
        json.dump(data, f, separators=(',', ':'))


if __name__ == '__main__':
    main()
