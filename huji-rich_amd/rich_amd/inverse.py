"""Sedov-Taylor inverse problem on a static jittered mesh: the workload of
the adjoint.

Parameters: the per-cell initial thermal-energy density u (N,) and the
blast energy E.  Initial cells: density 1, velocity 0, pressure
(gamma - 1) (u + E w), where w = 1 / V_hot on the cells whose mesh point lies
within `radius` of the origin (the blast energy spread uniformly as thermal
energy) and 0 elsewhere; u starts at the ambient p_amb / (gamma - 1).
Loss: sum over the `n_ring` cells nearest the circle r = r_ring of
(density - target)^2 after `steps` Heun steps, the target being the forward
run with E = E_target."""
import numpy as np
import torch

from . import mesh as rmesh
from .autograd import advance
from .gpu import RichGPU


class SedovInverse:
    def __init__(self, n=256, steps=400, every=20, gamma=5.0 / 3.0, p_amb=1e-5, radius=0.02,
                 r_ring=0.3, n_ring=1024, E_target=1.3, seed=11):
        self.mesh = rmesh.lattice(n, n, (-1.0, 1.0, -1.0, 1.0), jitter=1e-3, seed=seed,
                                  bc=rmesh.RIGID)
        m = self.mesh
        N = m.n_cells
        self.N, self.steps, self.every, self.gamma = N, steps, every, gamma
        r = np.hypot(m.pt_x[:N], m.pt_y[:N])
        hot = r < radius
        if not hot.any():
            hot = r <= r.min()
        w = np.zeros(N)
        w[hot] = 1.0 / m.volume[hot].sum()
        self.w = torch.from_numpy(w)
        self.u0 = torch.full((N,), p_amb / (gamma - 1.0), dtype=torch.float64)
        self.ring = torch.from_numpy(np.argsort(np.abs(r - r_ring), kind="stable")[:n_ring])
        self.solver = RichGPU(m, gamma=gamma, order=2, cfl=0.3)
        with torch.no_grad():
            self.target = self.forward(self.u0, torch.tensor(E_target, dtype=torch.float64))[0]

    def cells(self, u, E):
        p = (self.gamma - 1.0) * (u + E * self.w)
        one = torch.ones_like(u)
        zero = torch.zeros_like(u)
        return torch.stack([one, p, zero, zero])

    def forward(self, u, E):
        """Final cells (4, N) and the ring densities."""
        out = advance(self.cells(u, E), self.solver, self.steps, self.every)
        return out[0, self.ring], out

    def loss(self, u, E):
        rho, _ = self.forward(u, E)
        return torch.sum((rho - self.target) ** 2)

    def gradient(self, u, E):
        """(loss, dL/du (N,), dL/dE) by the device adjoint."""
        u = u.clone().requires_grad_(True)
        E = E.clone().requires_grad_(True)
        L = self.loss(u, E)
        L.backward()
        return L.item(), u.grad.detach(), E.grad.item()

    def loss_value(self, u, E):
        with torch.no_grad():
            return self.loss(u, E).item()

    def close(self):
        self.solver.close()
