"""The reference's ctypes declarations (tests/golden/ctypes_table.json, read from
napmo/__init__.py:95-585) applied to a private handle of libnapmo_gpu.so: the drop-in
check that the reference's own binding code binds this library unchanged."""
import ctypes
import json
import os

import numpy as np
import numpy.ctypeslib as npct

from tests import systems

TABLE = os.path.join(systems.ROOT, 'tests', 'golden', 'ctypes_table.json')
LIBPATH = os.path.join(systems.ROOT, 'napmo_amd', 'libnapmo_gpu.so')

# napmo/__init__.py:96-102
TYPES = {
    'None': None, 'c_void_p': ctypes.c_void_p, 'c_int': ctypes.c_int, 'c_long': ctypes.c_long,
    'c_double': ctypes.c_double, 'c_bool': ctypes.c_bool,
    'POINTER(c_double)': ctypes.POINTER(ctypes.c_double), 'POINTER(c_int)': ctypes.POINTER(ctypes.c_int),
    'a1df': npct.ndpointer(dtype=np.double, ndim=1, flags='CONTIGUOUS'),
    'a2df': npct.ndpointer(dtype=np.double, ndim=2, flags='CONTIGUOUS'),
    'a3df': npct.ndpointer(dtype=np.double, ndim=3, flags='CONTIGUOUS'),
    'a1di': npct.ndpointer(dtype=np.int32, ndim=1, flags='CONTIGUOUS'),
    'a1dli': npct.ndpointer(dtype=np.int64, ndim=1, flags='CONTIGUOUS'),
    'aptr': npct.ndpointer(ctypes.c_void_p, flags='C_CONTIGUOUS'),
}


def table():
    with open(TABLE) as f:
        return json.load(f)['symbols']


def grid_path_symbols():
    return sorted(k for k, e in table().items() if e['grid_path'])


class Bound:
    """Attribute access to libnapmo_gpu.so with exactly the reference's declared types."""

    def __init__(self):
        lib = ctypes.CDLL(LIBPATH)
        self.calls = set()
        for name, ent in table().items():
            if not ent['grid_path']:
                continue
            fn = getattr(lib, name)
            fn.restype = TYPES[ent['restype']]
            if ent['argtypes'] is not None:
                fn.argtypes = [TYPES[a] for a in ent['argtypes']]
            setattr(self, '_' + name, fn)

    def __getattr__(self, name):
        fn = self.__dict__.get('_' + name)
        if fn is None:
            raise AttributeError(name)
        self.calls.add(name)
        return fn
