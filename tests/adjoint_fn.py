"""torch.autograd wrapper of the C-ABI trajectory adjoint -- test harness only.

TrajectorySnapshots.apply(ctx, x0, v0, m, dt, nsteps, ckpt, snap_steps) returns
the positions at the snapshot steps (nsnap, n, 3); its backward calls
aster_traj_backward, so gradients of any loss of the snapshots with respect to
x0, v0 and m come from one hand-written adjoint pass."""
import numpy as np
import torch


class TrajectorySnapshots(torch.autograd.Function):
    @staticmethod
    def forward(fctx, actx, x0, v0, m, dt, nsteps, ckpt, snap_steps):
        actx.set_state(x0.detach().cpu().numpy(), v0.detach().cpu().numpy(), m.detach().cpu().numpy())
        snaps = actx.traj_forward(dt, nsteps, ckpt, snap_steps)
        fctx.actx = actx
        return torch.from_numpy(snaps).to(x0.dtype)

    @staticmethod
    def backward(fctx, grad_snaps):
        gx, gv, gm = fctx.actx.traj_backward(np.ascontiguousarray(grad_snaps.detach().cpu().numpy()))
        return None, torch.from_numpy(gx), torch.from_numpy(gv), torch.from_numpy(gm), None, None, None, None
