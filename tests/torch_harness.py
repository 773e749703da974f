"""torch.autograd.Function over the HIP adjoint (test harness only).

Only CPU tensors cross this boundary; the library owns device memory, so the
process needs no second GPU runtime.  Forward returns the per-pose energies;
backward calls the adjoint with the incoming per-pose gradient as the weight
vector and returns the DOF gradient and, when the table tensor requires grad,
the {dE/drc, dE/dac, dE/demin} gradient of every table entry.
"""
import numpy as np
import torch


class ScoreEnergy(torch.autograd.Function):
    @staticmethod
    def forward(ctx, dofs, table, hip_ctx, handles, T):
        d = dofs.detach().cpu().numpy().astype(np.float32)
        out = hip_ctx.adjoint(handles, d)
        ctx.hip_ctx, ctx.handles, ctx.T = hip_ctx, handles, T
        ctx.save_for_backward(dofs, table)
        return torch.from_numpy(out[:, 0].astype(np.float64))

    @staticmethod
    def backward(ctx, grad_e):
        dofs, table = ctx.saved_tensors
        d = dofs.detach().cpu().numpy().astype(np.float32)
        w = grad_e.detach().cpu().numpy().astype(np.float32)
        want_p = ctx.needs_input_grad[1]
        res = ctx.hip_ctx.adjoint(ctx.handles, d, w=w, pgrad=want_p, T=ctx.T)
        out, pg = (res if want_p else (res, None))
        width = dofs.shape[1]
        g = np.zeros((d.shape[0], width), np.float64)
        g[:, :7 - 1] = out[:, 1:7]
        if width == 26:
            g[:, 6:26] = out[:, 7:27]
        gp = torch.from_numpy(pg.astype(np.float64)) if want_p else None
        return torch.from_numpy(g), gp, None, None, None


def energy(dofs, table, hip_ctx, handles, T):
    return ScoreEnergy.apply(dofs, table, hip_ctx, handles, T)
