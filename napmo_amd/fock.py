"""Device-resident grid Fock build: the new fast entry point of the grid hot path.

``GridFock(grid, basis, lmax, precision)`` collocates Phi once per geometry (on the
GPU, point-major, padded to the MFMA tile) and builds the Poisson plan; each
``build(P)`` then runs rho = diag(Phi P Phi^T), the Becke-Dickson Poisson solve,
F = Phi^T diag(w bw v) Phi and the FP64 trace E = 1/2 sum P o F, all on the GPU.
This is the contraction of src/wavefunction/nwavefunction.cpp:77-107 applied to the
potential of napmo/wavefunction/ntwobody.py:15-46 with the density of
napmo/wavefunction/psi_hybrid.py:144-154.
"""
import ctypes

import numpy as np

from napmo_amd.system import cext

SCREEN_KEY = 11          # NAPMO_TUNE_INT8_SCREEN
DENSITY_SCREEN_KEY = 12  # NAPMO_TUNE_DENSITY_SCREEN
INT8_PACK_KEY = 13       # NAPMO_TUNE_INT8_PACK

PRECISIONS = {'fp64': 0, 'fp32': 1, 'bf16x3': 2, 'int8': 3}


class GridFock(object):

    def __init__(self, grid, basis, lmax, precision='fp64'):
        cext.require_gpu()
        self.grid = grid      # keeps the grid handle alive for the plan's lifetime
        self.basis = basis
        self.lmax = int(lmax)
        self.precision = precision
        plan = ctypes.c_void_p()
        cext.check(cext.napmo_fock_plan_new(grid._this, basis._this, self.lmax, PRECISIONS[precision],
                                            ctypes.byref(plan)))
        self._this = plan.value
        n, nb, nc, lm = ctypes.c_long(), ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        cext.check(cext.napmo_fock_plan_info(self._this, ctypes.byref(n), ctypes.byref(nb),
                                             ctypes.byref(nc), ctypes.byref(lm)))
        self.npoint, self.nbasis = n.value, nb.value

    def __del__(self):
        if getattr(self, '_this', None):
            cext.napmo_fock_plan_del(self._this)
            self._this = None

    def build(self, P, want_grid=False):
        """Returns (F [n x n], E_trace, E_grid[, rho, v]) for a host density matrix P."""
        P = np.ascontiguousarray(P, dtype=np.float64)
        n = self.nbasis
        assert P.shape == (n, n)
        F = np.zeros((n, n))
        E = np.zeros(2)
        rho = np.zeros(self.npoint) if want_grid else None
        v = np.zeros(self.npoint) if want_grid else None
        ptr = lambda a: None if a is None else a.ctypes.data  # noqa: E731
        cext.check(cext.napmo_grid_fock_build(self._this, ptr(P), ptr(F), ptr(E), ptr(rho), ptr(v)))
        if want_grid:
            return F, E[0], E[1], rho, v
        return F, E[0], E[1]

    def build_orbitals(self, C, eta=2.0, want_grid=False):
        """Orbital form (rho = eta sum_i (Phi C)_i^2, P = eta C C^T); C [n x nocc]."""
        C = np.ascontiguousarray(C, dtype=np.float64)
        n = self.nbasis
        assert C.ndim == 2 and C.shape[0] == n
        F = np.zeros((n, n))
        E = np.zeros(2)
        rho = np.zeros(self.npoint) if want_grid else None
        v = np.zeros(self.npoint) if want_grid else None
        ptr = lambda a: None if a is None else a.ctypes.data  # noqa: E731
        cext.check(cext.napmo_grid_fock_build_orbitals(self._this, C.ctypes.data, C.shape[1], eta,
                                                       ptr(F), ptr(E), ptr(rho), ptr(v)))
        if want_grid:
            return F, E[0], E[1], rho, v
        return F, E[0], E[1]

    def build_orbitals_device(self, dC, nocc, eta, dF, dE):
        cext.check(cext.napmo_grid_fock_build_orbitals_device(self._this, dC, nocc, eta, dF, dE))

    def build_device(self, dP, dF, dE):
        """Asynchronous build on device pointers (ints), on the plan's stream."""
        cext.check(cext.napmo_grid_fock_build_device(self._this, dP, dF, dE))

    def contract(self, v):
        """Phi^T diag(w bw v) Phi for a potential v [N] on the grid (any sign)."""
        F = np.zeros((self.nbasis, self.nbasis))
        cext.check(cext.napmo_fock_plan_contract(self._this, np.ascontiguousarray(v, dtype=np.float64), F))
        return F

    def overlap(self):
        """S = Phi^T diag(w bw) Phi from the resident Phi."""
        S = np.zeros((self.nbasis, self.nbasis))
        cext.check(cext.napmo_grid_overlap(self._this, S))
        return S

    def centres(self):
        """Basis centres [ncentre x 3] (distinct function origins, first-appearance order)."""
        nc = ctypes.c_int()
        cext.check(cext.napmo_fock_plan_centres(self._this, ctypes.byref(nc), None))
        xyz = np.zeros((nc.value, 3))
        cext.check(cext.napmo_fock_plan_centres(self._this, ctypes.byref(nc), xyz.ctypes.data))
        return xyz

    def energy_dP(self):
        """dE/dP [n x n] of the last build's E_grid = 1/2 rho^T W V rho, exact for the discrete
        Poisson operator V: Phi^T diag((W v + V^T (W rho)) / 2) Phi."""
        out = np.zeros((self.nbasis, self.nbasis))
        cext.check(cext.napmo_grid_fock_energy_dP(self._this, out.ctypes.data))
        return out

    def energy_dP_device(self, d_out):
        cext.check(cext.napmo_grid_fock_energy_dP_device(self._this, d_out))

    def poisson_adjoint(self, y):
        """V^T y [N] for the plan's grid and lmax (V the Becke-Poisson operator of the build)."""
        y = np.ascontiguousarray(y, dtype=np.float64)
        z = np.zeros(self.npoint)
        cext.check(cext.napmo_grid_poisson_adjoint(self._this, y.ctypes.data, z.ctypes.data))
        return z

    def gradient(self, P):
        """dE/dR_c [ncentre x 3] of the last build's E = 1/2 sum w bw rho v (same P),
        grid fixed, basis functions moving with their centres; exact for the discrete
        operator (the symmetrised weight (W v + V^T (W rho)) / 2)."""
        P = np.ascontiguousarray(P, dtype=np.float64)
        g = np.zeros((len(self.centres()), 3))
        cext.check(cext.napmo_grid_fock_gradient(self._this, P.ctypes.data, g.ctypes.data))
        return g

    def gradient_device(self, dP, dG):
        cext.check(cext.napmo_grid_fock_gradient_device(self._this, dP, dG))

    def set_screen(self, S):
        """int8 plans: skip, per chunk of points, the 16 x 16 blocks of F bounded by 2^-S in
        every element (0: off), so each element moves by at most (chunks) 2^-S."""
        cext.check(cext.napmo_fock_plan_set_tuning(self._this, SCREEN_KEY, int(S)))

    def set_density_screen(self, S):
        """Orbital-form density on the streamed kernel: skip, per 16-point tile, the
        16-function groups whose terms of X = Phi C are bounded by 2^-S (0: off), so each
        X_pc moves by at most (ld / 16) 2^-S."""
        cext.check(cext.napmo_fock_plan_set_tuning(self._this, DENSITY_SCREEN_KEY, int(S)))

    def set_int8_pack(self, on):
        """int8 plans: pack each chunk's kept blocks into as few workgroup tasks as fit (the
        same results bit for bit)."""
        cext.check(cext.napmo_fock_plan_set_tuning(self._this, INT8_PACK_KEY, int(bool(on))))

    def screen_stats(self):
        """int8 plans: (blocks x chunks computed by the last contraction, all of them)."""
        kept, total = ctypes.c_long(), ctypes.c_long()
        cext.check(cext.napmo_fock_plan_int8_screen_stats(self._this, ctypes.byref(kept), ctypes.byref(total)))
        return kept.value, total.value

    def profile(self, max_builds):
        cext.check(cext.napmo_fock_plan_profile(self._this, max_builds))

    def profile_read(self):
        """{phase name: average ms} over the builds recorded since ``profile``."""
        avg = np.zeros(16)
        nb = ctypes.c_int()
        cext.check(cext.napmo_fock_plan_profile_read(self._this, avg, 16, ctypes.byref(nb)))
        names = [cext.napmo_fock_phase_name(k).decode() for k in range(16)]
        return {nm: float(a) for nm, a in zip(names, avg) if nm}, nb.value

    @property
    def stream(self):
        return cext.napmo_fock_plan_stream(self._this)

    def synchronize(self):
        cext.check(cext.napmo_fock_plan_sync(self._this))
