"""torch.autograd.Function over the C-ABI's discrete adjoint, for gradient
checks only (the forward and adjoint run in libnebulasem_amd.so; torch only
composes them with pointwise maps and supplies the objective's gradient).

    q_N = RK3Steps.apply(q0, dg, nsteps, checkpoint)

q0, q_N: float64 tensors of shape (5, nb): rho', m_x, m_y, m_z, Theta'.
"""
import numpy as np
import torch


class RK3Steps(torch.autograd.Function):
    @staticmethod
    def forward(ctx, q0, dg, nsteps, checkpoint):
        q = q0.detach().cpu().numpy().astype(np.float64)
        dg.set_conserved(q[0], np.ascontiguousarray(q[1:4].T), q[4])
        dg.rk3_steps(nsteps, 64)
        rp, m, tp = dg.get_conserved()
        m = m.reshape(-1, 3)
        ctx.save_for_backward(q0)
        ctx.dg, ctx.nsteps, ctx.checkpoint = dg, nsteps, checkpoint
        out = np.stack([rp, m[:, 0], m[:, 1], m[:, 2], tp])
        return torch.from_numpy(out).to(q0.device)

    @staticmethod
    def backward(ctx, grad):
        (q0,) = ctx.saved_tensors
        q = q0.detach().cpu().numpy().astype(np.float64)
        ctx.dg.set_conserved(q[0], np.ascontiguousarray(q[1:4].T), q[4])
        lam = ctx.dg.rk3_adjoint(ctx.nsteps, grad.detach().cpu().numpy(), ctx.checkpoint)
        return torch.from_numpy(lam).to(q0.device), None, None, None
