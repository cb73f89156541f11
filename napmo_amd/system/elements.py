"""Atomic data used by the grids (radii in bohr), from napmo_amd/data/elements.json."""
import json
import os

_DB = None


def AtomicElementsDatabase():
    """Dict symbol -> {'atomic_number', 'atomic_radii', 'atomic_radii_2', 'mass', 'charge'}."""
    global _DB
    if _DB is None:
        with open(os.path.join(os.path.dirname(__file__), '..', 'data', 'elements.json')) as f:
            _DB = json.load(f)
    return _DB


def atomic_radius(symbol):
    """Radius choice of napmo/grids/radial.py:20-48."""
    db = AtomicElementsDatabase()
    radii = db.get('H', {}).get('atomic_radii_2', 1.0)
    k = symbol.find('_')
    if k > 0:
        radii = db.get(symbol[:k], {}).get('atomic_radii', 1.0)
    if symbol in db:
        radii = db[symbol].get('atomic_radii', 1.0)
    return radii
