"""torch.autograd glue for the grid Coulomb energy E(P, R) of the device Fock build.

E = 1/2 sum_p w_p bw_p rho_p v_p with rho = diag(Phi(R) P Phi(R)^T) and v the Becke-Poisson
potential of rho, on a fixed Becke grid; R are the centres the basis functions move with.
The backward is hand-written on the GPU:

* with W = w bw and V the discrete Becke-Poisson operator, E = 1/2 rho^T W V rho, so
  dE/drho = ssym = (W v + V^T (W rho)) / 2, V^T the exactly transposed projection, radial
  solve, synthesis tables and synthesis (napmo_grid_poisson_adjoint);
* dE/dP = Phi^T diag(ssym) Phi (napmo_grid_fock_energy_dP). The build's own F uses W v alone
  and equals dE/dP only as far as V is self-adjoint in the W inner product;
* dE/dR_c = -2 sum_p ssym_p sum_{mu on c} grad phi_mu (Phi P)_p mu, with grad Phi
  recomputed in the kernel rather than stored (napmo_grid_fock_gradient).

Tensors may live on the GPU (device pointers go straight to the library) or on the CPU.
torch is used only for this glue; the arithmetic is in libnapmo_gpu.so.
"""
import numpy as np
import torch

from napmo_amd.fock import GridFock


class GridEnergyModel(object):
    """A fixed grid, atom symbols and basis shells per atom; the device plan (Phi
    collocated at the current centres) is rebuilt whenever R changes."""

    def __init__(self, grid, symbols, shells_per_atom, lmax, precision='fp64'):
        self.grid = grid
        self.symbols = list(symbols)
        self.shells = shells_per_atom
        self.lmax = int(lmax)
        self.precision = precision
        self._R = None
        self._plan = None
        self._atom_of = None

    def plan(self, R):
        from napmo_amd.system.species import make_basis
        R = np.ascontiguousarray(R, dtype=np.float64).reshape(-1, 3)
        if self._R is None or not np.array_equal(R, self._R):
            atoms = [(s, tuple(float(c) for c in r)) for s, r in zip(self.symbols, R)]
            self._basis = make_basis(atoms, shells_per_atom=self.shells)
            self._plan = None  # release the previous plan's device memory first
            self._plan = GridFock(self.grid, self._basis, self.lmax, self.precision)
            self._R = R.copy()
            cen = self._plan.centres()
            self._atom_of = []
            for c in cen:
                hit = np.flatnonzero(np.all(R == c, axis=1))
                if len(hit) != 1:
                    raise ValueError('basis centre %s is not exactly one atom position' % (c,))
                self._atom_of.append(int(hit[0]))
        return self._plan

    @property
    def nbasis(self):
        return self._plan.nbasis if self._plan else None


def _dev(t):
    return t.is_cuda


class GridEnergy(torch.autograd.Function):
    """E = GridEnergy.apply(P, R, model): P [n x n] float64, R [natom x 3] float64."""

    @staticmethod
    def forward(ctx, P, R, model):
        plan = model.plan(R.detach().cpu().numpy())
        P = P.detach().contiguous()
        if _dev(P):
            F = torch.empty_like(P)
            E = torch.empty(2, dtype=torch.float64, device=P.device)
            torch.cuda.synchronize(P.device)
            plan.build_device(P.data_ptr(), F.data_ptr(), E.data_ptr())
            plan.synchronize()
            e_grid = E[1].clone()
        else:
            Fn, _, e = plan.build(P.numpy())
            F = torch.from_numpy(Fn)
            e_grid = torch.tensor(e, dtype=torch.float64)
        ctx.model = model
        ctx.save_for_backward(P, R)
        return e_grid

    @staticmethod
    def backward(ctx, gE):
        P, R = ctx.saved_tensors
        model = ctx.model
        gP = gR = None
        plan = model.plan(R.detach().cpu().numpy())
        # rebuild with this P so the plan's density and potential are the ones E was taken from
        if _dev(P):
            Fd = torch.empty_like(P)
            E = torch.empty(2, dtype=torch.float64, device=P.device)
            torch.cuda.synchronize(P.device)
            plan.build_device(P.data_ptr(), Fd.data_ptr(), E.data_ptr())
            if ctx.needs_input_grad[0]:
                dP = torch.empty_like(P)
                plan.energy_dP_device(dP.data_ptr())
                plan.synchronize()
                gP = gE * dP
            if ctx.needs_input_grad[1]:
                G = torch.empty((len(model._atom_of), 3), dtype=torch.float64, device=P.device)
                plan.gradient_device(P.data_ptr(), G.data_ptr())
                plan.synchronize()
                gc = G.cpu().numpy()
        else:
            plan.build(P.numpy())
            if ctx.needs_input_grad[0]:
                gP = gE * torch.from_numpy(plan.energy_dP())
            if ctx.needs_input_grad[1]:
                gc = plan.gradient(P.numpy())
        if ctx.needs_input_grad[1]:
            g = np.zeros((R.shape[0], 3))
            for c, a in enumerate(model._atom_of):
                g[a] += gc[c]
            gR = gE * torch.as_tensor(g, dtype=R.dtype, device=R.device)
        return gP, gR, None


def grid_energy(P, R, model):
    return GridEnergy.apply(P, R, model)


__all__ = ['GridEnergyModel', 'GridEnergy', 'grid_energy']
