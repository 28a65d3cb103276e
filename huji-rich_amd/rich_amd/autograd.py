"""torch.autograd bridge to the device adjoint: a thin harness that lets a
torch loss differentiate through device-resident RICH steps.

`advance(cells0, solver, nsteps, every)` runs `nsteps` Heun steps of the
RichGPU context `solver` from the initial cells (4, N) float64 tensor and
returns the final cells; its backward is rg_adjoint_backward (checkpoints
every `every` steps, each segment recomputed once).  A context records one
forward run at a time: a backward whose run was replaced by a later forward
on the same context raises instead of returning the later run's gradient."""
import numpy as np
import torch


class RichAdvance(torch.autograd.Function):
    @staticmethod
    def forward(ctx, cells0, solver, nsteps, every):
        c = np.ascontiguousarray(cells0.detach().cpu().numpy(), dtype=np.float64)
        solver.set_cells(c)
        solver.init_extensives()
        solver.adjoint_forward(int(nsteps), int(every), record_dt=False)
        ctx.solver = solver
        ctx.generation = solver.adjoint_generation
        return torch.from_numpy(solver.get_cells()).to(cells0.device)

    @staticmethod
    def backward(ctx, grad):
        if ctx.solver.adjoint_generation != ctx.generation:
            raise RuntimeError("the solver ran another forward after this one: its device "
                               "checkpoints belong to that run, so this graph cannot be "
                               "differentiated (use one solver per pending backward)")
        g0 = ctx.solver.adjoint_backward(grad.detach().cpu().numpy().astype(np.float64))
        return torch.from_numpy(g0).to(grad.device), None, None, None


def advance(cells0, solver, nsteps, every=20):
    return RichAdvance.apply(cells0, solver, nsteps, every)
