"""Batched grid Fock builds of many independent molecules, optionally replayed from a
HIP graph (napmo_batch_* in libnapmo_gpu.so).

``BatchFock(plans, nlanes)`` takes GridFock plans (one per molecule, each with its own
grid and basis) and spreads their builds over ``nlanes`` HIP streams; ``nlanes=0`` runs
every phase of all the builds as one launch (FP64 plans). ``capture`` records
one build of every molecule into a graph on fixed device buffers; ``replay`` re-runs it,
one graph launch per SCF iteration, after the caller has written the next density
matrices into the same buffers.
"""
import ctypes

from napmo_amd.system import cext


class BatchFock(object):

    def __init__(self, plans, nlanes=4):
        self.plans = list(plans)   # the batch borrows the plans: keep them alive
        m = len(self.plans)
        arr = (ctypes.c_void_p * m)(*[p._this for p in self.plans])
        out = ctypes.c_void_p()
        cext.check(cext.napmo_batch_plan_new(m, arr, int(nlanes), ctypes.byref(out)))
        self._this = out.value

    def __del__(self):
        if getattr(self, '_this', None):
            cext.napmo_batch_plan_del(self._this)
            self._this = None

    @staticmethod
    def _ptrs(values):
        return (ctypes.c_void_p * len(values))(*values)

    def build(self, dPs, dFs, dEs):
        """Asynchronous batch build on device pointers (lists of ints, one per molecule)."""
        cext.check(cext.napmo_batch_build(self._this, self._ptrs(dPs), self._ptrs(dFs), self._ptrs(dEs)))

    def build_orbitals(self, dCs, noccs, eta, dFs, dEs):
        """Orbital form (streams): dCs [n_m x nocc_m] device pointers, P_m = eta C C^T."""
        nocc = (ctypes.c_int * len(noccs))(*[int(k) for k in noccs])
        cext.check(cext.napmo_batch_build_orbitals(self._this, self._ptrs(dCs), nocc, float(eta),
                                                   self._ptrs(dFs), self._ptrs(dEs)))

    def capture_orbitals(self, dCs, noccs, eta, dFs, dEs):
        nocc = (ctypes.c_int * len(noccs))(*[int(k) for k in noccs])
        cext.check(cext.napmo_batch_capture_orbitals(self._this, self._ptrs(dCs), nocc, float(eta),
                                                     self._ptrs(dFs), self._ptrs(dEs)))

    def capture(self, dPs, dFs, dEs):
        cext.check(cext.napmo_batch_capture(self._this, self._ptrs(dPs), self._ptrs(dFs), self._ptrs(dEs)))

    def replay(self):
        cext.check(cext.napmo_batch_replay(self._this))

    def synchronize(self):
        cext.check(cext.napmo_batch_sync(self._this))
