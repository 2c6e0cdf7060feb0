"""torch.autograd binding of the fixed-step RK4 solve and its discrete adjoint.

The reference differentiates a solve through ChainRulesCore rules (SciMLSensitivity); here the
rule is a ``torch.autograd.Function``: the forward pass runs the device RK4 loop keeping
checkpoints, the backward pass runs the discrete RK4 adjoint (segments recomputed from the
checkpoints, four generated vector-Jacobian products per step).  Gradients flow to the initial
state and to grid-field tensors (e.g. a learnable per-cell diffusivity).
"""
from __future__ import annotations

from typing import Dict, Optional

import numpy as np

from .problem import ODEProblem, _torch


def _function():
    torch = _torch()

    class _RK4Adjoint(torch.autograd.Function):
        @staticmethod
        def forward(ctx, u0, prob, t0, dt, nsteps, ckpt_every, names, *fields):
            for name, fld in zip(names, fields):
                prob.set_grid_field(name, fld.detach())
            u = u0.detach().clone().contiguous()
            prob.forward_checkpointed(u, t0, dt, nsteps, ckpt_every)
            ctx.prob = prob
            ctx.names = names
            ctx.dtypes = [f.dtype for f in fields]
            return u

        @staticmethod
        def backward(ctx, gu):
            prob = ctx.prob
            lam = gu.detach().to(prob.torch_dtype()).clone().contiguous()
            ng = int(np.prod(prob.lowered.n))
            naux = len(prob.lowered.aux)
            g = torch.zeros(max(1, naux) * ng, dtype=torch.float64, device=lam.device)
            prob.adjoint(lam, g if naux else None)
            grads = []
            for name, dt_ in zip(ctx.names, ctx.dtypes):
                a = prob.lowered.aux.index(name)
                grads.append(g[a * ng:(a + 1) * ng].to(dt_))
            return (lam, None, None, None, None, None, None, *grads)

    return _RK4Adjoint


def rk4_solve(prob: ODEProblem, u0, t0: float, dt: float, nsteps: int, ckpt_every: Optional[int] = 20,
              grid_fields: Optional[Dict[str, object]] = None):
    """Differentiable RK4 solve: returns u(t0 + nsteps*dt).  ``grid_fields`` maps grid-field
    names to contiguous CUDA tensors (flat, first axis fastest) that are bound for the solve and
    receive gradients.  ``ckpt_every``: steps between checkpoints, 0 to keep every stage state,
    None to choose by free device memory.  The problem keeps one checkpoint set: backward must
    follow its own forward before the next solve of the same problem."""
    names = tuple(grid_fields.keys()) if grid_fields else ()
    fields = tuple(grid_fields[k] for k in names) if grid_fields else ()
    if ckpt_every is None:
        ckpt_every = prob.checkpoint_interval(nsteps)
    return _function().apply(u0, prob, float(t0), float(dt), int(nsteps), int(ckpt_every), names, *fields)
