"""Forward and backward runs of the differentiable form without torch.

DiffRun drives libpg's pg_adjoint_forward / pg_adjoint_backward for a family
of setups parametrised by theta = (bulk modulus K, contact-stiffness scale,
friction coefficient).  The parameters enter the stepper through
  c_g = K * alpha_g             every grain's micromodulus (proportional to K)
  Kn  = scale * K * alpha_mean  normal contact stiffness (node-node and wall)
  mu                            friction coefficient
so with the library's raw derivatives g = (sum_g c_g dL/dc_g, dL/dKn, dL/dmu):
  dL/dK = (g0 + Kn g1) / K,  dL/dscale = Kn g1 / scale,  dL/dmu = g2.
The critical stretch also depends on K, but bond status is frozen to the
forward pass, so it carries no derivative.  perigrain_amd.adjoint wraps this
in a torch.autograd.Function.
"""
import numpy as np

from . import setups
from .stepper import Stepper


def trajectory_bytes(s, steps):
    """Device bytes that keeping every step's state of a recorded pass takes:
    per step the node records (X and u), velocities, grain boxes and walls."""
    ns = (s.n_nodes + 1 + 63) // 64 * 64  # nodes and the ghost record, padded
    return steps * (ns * 3 * s.dim * 8 + s.n_grains * 2 * s.dim * 8 + 6 * 8)


def auto_checkpoint_every(s, steps, budget=16 << 30):
    """checkpoint_every for pg_adjoint_forward: the whole run, so that the
    backward pass reuses every step's state instead of recomputing it, when
    that state fits in `budget` bytes of device memory; otherwise
    checkpoints every ceil(sqrt(steps)) steps (each segment recomputed once
    in the backward pass)."""
    if trajectory_bytes(s, steps) <= budget:
        return int(steps)
    return int(np.ceil(np.sqrt(steps)))


class DiffRun:
    """Forward and backward runs of one problem family on a device.
    checkpoint_every None: auto_checkpoint_every."""

    def __init__(self, make_setup, steps, checkpoint_every=None, device=0):
        self.make_setup = make_setup  # (K, scale, mu) -> (Setup, Params)
        self.steps = int(steps)
        self.checkpoint_every = None if checkpoint_every is None else int(checkpoint_every)
        self.device = device

    def forward(self, K, scale, mu):
        s, p = self.make_setup(K, scale, mu)
        ck = self.checkpoint_every or auto_checkpoint_every(s, self.steps)
        g = Stepper(s, p, device=self.device)
        hist = g.adjoint_forward(self.steps, ck)
        return hist, (g, K, scale, mu, s.normal_stiffness)

    def plain(self, K, scale, mu):
        """Reaction history of the same run with ordinary steps (no
        recording), one pg_step per step."""
        s, p = self.make_setup(K, scale, mu)
        ns = 4 if s.dim == 2 else 6
        out = np.zeros((self.steps, ns, s.dim))
        with Stepper(s, p, device=self.device) as g:
            for t in range(self.steps):
                g.step(1)
                out[t] = g.wall()[1]
        return out

    def backward(self, state, reaction_bar):
        g, K, scale, mu, Kn = state
        try:
            raw = g.adjoint_backward(reaction_bar)
        finally:
            g.close()
        dK = (raw[0] + Kn * raw[1]) / K
        ds = Kn * raw[1] / scale
        return np.array([dK, ds, raw[2]]), raw


def top_wall_misfit(hist, target, side=2):
    """L2 misfit of the top-wall reaction history (side 2 in 2D: top) and
    its cotangent with respect to the history."""
    d = hist[:, side, :] - target[:, side, :]
    bar = np.zeros_like(hist)
    bar[:, side, :] = 2.0 * d
    return float((d ** 2).sum()), bar


def compression_2d_run(steps=2000, checkpoint_every=None, device=0, **kw):
    """The differentiable 2D compression (64 disks and squares, ~250 nodes
    each, top wall at -0.5 m/s, dt = 2e-8) as a DiffRun."""
    def make(K, scale, mu):
        return setups.compression_2d_diff(K=K, contact_scale=scale, mu=mu, **kw)
    return DiffRun(make, steps, checkpoint_every, device)
