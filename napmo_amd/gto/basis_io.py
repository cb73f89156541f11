"""Reader for deMon2k-format basis-set files (EMSL Basis Set Exchange export).

Follows the format the reference reads in ``napmo/gto/basis_set.py:56-120``: a block
starts at a line ``O-<NAME> <SYMBOL> (...)``; after comments, the next line holds
the number of contractions; each contraction is a line ``<index> <l> <nprim>``
followed by ``nprim`` lines ``<exponent> <coefficient>``.
"""
import re


def parse_demon2k(text, symbol):
    """Return ``[{'l': int, 'prim': [...], 'cont': [...]}, ...]`` for ``symbol``.

    Raises ValueError when the file holds no block for ``symbol`` (the reference
    raises the same error type, ``napmo/gto/basis_set.py:90-96``).
    """
    lines = [ln.split('#', 1)[0].rstrip() for ln in text.splitlines()]
    lines = [ln for ln in lines if ln.strip()]
    pat = re.compile(r'^O-\S*\s+' + re.escape(symbol) + r'(\s|$)')
    start = next((i for i, ln in enumerate(lines) if pat.match(ln)), None)
    if start is None:
        raise ValueError('Basis for particle ' + symbol + ' not found!')
    ncont = int(lines[start + 1].split()[0])
    pos = start + 2
    shells = []
    for _ in range(ncont):
        head = lines[pos].split()
        l, nprim = int(head[1]), int(head[2])
        pos += 1
        prim, cont = [], []
        for _ in range(nprim):
            e, c = lines[pos].split()[:2]
            prim.append(float(e.replace('D', 'E')))
            cont.append(float(c.replace('D', 'E')))
            pos += 1
        shells.append({'l': l, 'prim': prim, 'cont': cont})
    return shells
