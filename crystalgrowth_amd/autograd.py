"""torch.autograd binding of the adjoint (kg_adjoint).

`kobayashi_loss(params, T0, ...)` returns L = sum (phi_N - target)^2 of an
N-step run as a torch scalar whose backward gives dL/d(params) and dL/dT0
from the hand-written reverse kernels (one forward + reverse pass on the GPU;
the tensors themselves live on the host).
"""
from __future__ import annotations

import numpy as np
import torch

from .kobayashi import Solver, default_params

PARAM_NAMES = ("tau", "eps_bar", "K", "delta", "theta0", "alpha", "gamma", "teq", "noise_a")


class _KobayashiLoss(torch.autograd.Function):
    @staticmethod
    def forward(ctx, p, T0, names, base, phi0, theta, target, nsteps, every, precision, step0):
        vals = {n: float(v) for n, v in zip(names, p.detach().cpu().double().tolist())}
        prm = base.replace(**vals)
        ny, nx = target.shape
        with Solver(nx, ny, precision, prm) as s:
            s.set_fields(phi0, T0.detach().cpu().numpy(), theta)
            s.step_count = step0
            g, dT, _ = s.adjoint(nsteps, target, every, want_T0=True)
        gp = torch.tensor([g["d_" + n] for n in names], dtype=p.dtype)
        ctx.save_for_backward(gp, torch.from_numpy(dT.astype(np.float64)).to(T0.dtype))
        return torch.tensor(g["loss"], dtype=torch.float64)

    @staticmethod
    def backward(ctx, grad):
        gp, gT = ctx.saved_tensors
        return (grad * gp, grad * gT) + (None,) * 9


def kobayashi_loss(p: torch.Tensor, T0: torch.Tensor, names=("delta", "K", "eps_bar", "tau"),
                   base=None, phi0=None, theta=None, target=None, nsteps: int = 100,
                   every: int = 50, precision: str = "f32", step0: int = 0) -> torch.Tensor:
    """Differentiable loss of an nsteps run.

    p      -- values of the parameters `names` (1-D tensor)
    T0     -- initial temperature field (ny, nx)
    base   -- kg_params for everything else (reference defaults if None)
    phi0, theta -- initial phase and angle fields (numpy); target -- phi to match
    """
    if base is None:
        base = default_params()
    names = tuple(names)
    for n in names:
        if n not in PARAM_NAMES:
            raise ValueError(f"no adjoint for parameter {n!r}")
    return _KobayashiLoss.apply(p, T0, names, base, phi0, theta, target, nsteps, every,
                                precision, step0)
