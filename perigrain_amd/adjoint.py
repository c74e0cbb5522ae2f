"""Differentiable form behind torch: the wall-reaction history of a run as a
function of theta = (bulk modulus K, contact-stiffness scale, friction
coefficient), with libpg's hand-written adjoint (pg_adjoint_forward /
pg_adjoint_backward, driven by perigrain_amd.diffrun) as the backward of a
torch.autograd.Function.

torch is imported by this module; import it (or this module) before any
other module that loads libpg, so that one HIP runtime serves both.
"""
import torch

from .diffrun import DiffRun, compression_2d_run  # noqa: F401


class WallReactionHistory(torch.autograd.Function):
    """theta = (K, contact-stiffness scale, mu) -> reaction history
    [steps, sides, dim] (float64, CPU)."""

    @staticmethod
    def forward(ctx, theta, run):
        K, scale, mu = (float(x) for x in theta.detach().cpu().tolist())
        hist, state = run.forward(K, scale, mu)
        ctx.run = run
        ctx.state = state
        return torch.from_numpy(hist)

    @staticmethod
    def backward(ctx, grad_hist):
        g, _ = ctx.run.backward(ctx.state, grad_hist.detach().cpu().numpy())
        return torch.tensor(g, dtype=torch.float64), None


def reaction_history(theta, run):
    return WallReactionHistory.apply(theta, run)


def top_wall_misfit(hist, target, side=2):
    """L2 misfit of the top-wall reaction history (side 2 in 2D: top)."""
    return ((hist[:, side, :] - target[:, side, :]) ** 2).sum()


