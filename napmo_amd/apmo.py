"""Multi-species (APMO) grid Coulomb build on the GPU.

``APMOFock([(grid, basis, charge, lmax), ...], precision)`` builds one device plan per
species (Phi collocated once, Poisson tables, cross-species synthesis lists). Each
``build([P_1, P_2, ...])`` returns, per species s,

    F_s = Phi_s^T diag(w bw v_eff_s) Phi_s,  v_eff_s = q_s^2 V[rho_s] + q_s sum_t q_t V[rho_t]

which is q_s^2 times the species' Coulomb matrix (napmo/wavefunction/psi_numeric.py:152-169)
plus its coupling matrix (psi_numeric.py:217-241, psi_hybrid.py:81-110), and the energies
(1/2 sum P o F, 1/2 sum w bw rho v_eff, own-Coulomb part, coupling part). Species may
share a grid (the reference's arrangement, napmo/solver/hf_solver.py:95-98) or each have
their own, in which case the other species' expansions are synthesised at its points.
"""
import ctypes

import numpy as np

from napmo_amd.fock import PRECISIONS
from napmo_amd.system import cext


class APMOFock(object):

    def __init__(self, species, precision='fp64', one_launch=False):
        cext.require_gpu()
        self.species = list(species)   # keeps the grid / basis handles alive
        ns = len(self.species)
        grids = (ctypes.c_void_p * ns)(*[g._this for g, _, _, _ in self.species])
        bases = (ctypes.c_void_p * ns)(*[b._this for _, b, _, _ in self.species])
        charges = (ctypes.c_double * ns)(*[float(q) for _, _, q, _ in self.species])
        lmax = (ctypes.c_int * ns)(*[int(l) for _, _, _, l in self.species])
        plan = ctypes.c_void_p()
        cext.check(cext.napmo_apmo_plan_new(ns, grids, bases, charges, lmax, PRECISIONS[precision],
                                            ctypes.byref(plan)))
        self._this = plan.value
        self.nbasis = [b.get('nbasis') for _, b, _, _ in self.species]
        if one_launch:
            self.set_one_launch(True)

    def set_one_launch(self, on):
        """Every phase of a build as one launch for all species (FP64 plans)."""
        cext.check(cext.napmo_apmo_plan_set_one_launch(self._this, int(bool(on))))

    def set_tuning(self, species, key, value):
        """A ``NAPMO_TUNE_*`` key of one species' Fock plan (not in one-launch mode)."""
        cext.check(cext.napmo_apmo_plan_set_tuning(self._this, int(species), int(key), int(value)))

    def set_screen(self, S):
        """int8 plans: the block screening at 2^-S of every species' contraction (0: off)."""
        from napmo_amd.fock import SCREEN_KEY
        for s in range(len(self.species)):
            self.set_tuning(s, SCREEN_KEY, S)

    def __del__(self):
        if getattr(self, '_this', None):
            cext.napmo_apmo_plan_del(self._this)
            self._this = None

    def build(self, Ps):
        """Ps: one [n_s x n_s] density matrix per species. Returns (list of F_s,
        E [nspecies x 4])."""
        ns = len(self.species)
        Ps = [np.ascontiguousarray(P, dtype=np.float64) for P in Ps]
        for P, n in zip(Ps, self.nbasis):
            assert P.shape == (n, n)
        Fs = [np.zeros((n, n)) for n in self.nbasis]
        E = np.zeros(4 * ns)
        pP = (ctypes.c_void_p * ns)(*[P.ctypes.data for P in Ps])
        pF = (ctypes.c_void_p * ns)(*[F.ctypes.data for F in Fs])
        cext.check(cext.napmo_apmo_build(self._this, pP, pF, E.ctypes.data))
        return Fs, E.reshape(ns, 4)

    def build_device(self, dPs, dFs, dE):
        """Asynchronous build on device pointers (lists of ints); see ``synchronize``."""
        ns = len(self.species)
        pP = (ctypes.c_void_p * ns)(*dPs)
        pF = (ctypes.c_void_p * ns)(*dFs)
        cext.check(cext.napmo_apmo_build_device(self._this, pP, pF, dE))

    def capture(self, dPs, dFs, dE):
        """One-launch mode: record one build on these device pointers into a HIP graph."""
        ns = len(self.species)
        pP = (ctypes.c_void_p * ns)(*dPs)
        pF = (ctypes.c_void_p * ns)(*dFs)
        cext.check(cext.napmo_apmo_capture(self._this, pP, pF, dE))

    def replay(self):
        cext.check(cext.napmo_apmo_replay(self._this))

    def synchronize(self):
        cext.check(cext.napmo_apmo_sync(self._this))
