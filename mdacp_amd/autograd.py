"""torch.autograd.Function over the hand-written adjoint of libmdx.

    loss = LJTrajectoryLoss.apply(q0, eps, sigma, p0, L, opts)

q0 (N, 3), eps and sigma (0-d) are differentiable; the forward runs the
trajectory on the GPU through mdx_lj_trajectory and the gradients come from
its adjoint (Hessian-vector products of the forces, dF/deps, dF/dsigma, with
recomputation from checkpoints), not from torch's autograd.  Tensors may
live on any device; they are staged through host memory.
"""
from __future__ import annotations

import numpy as np
import torch

from .mdx import System


class LJTrajectoryLoss(torch.autograd.Function):
    @staticmethod
    def forward(ctx, q0, eps, sigma, p0, L, opts=None):
        opts = dict(opts or {})
        dt = opts.pop("dt", 0.001)
        qn = q0.detach().to("cpu", torch.float64).numpy()
        pn = p0.detach().to("cpu", torch.float64).numpy()
        with System(np.asarray(L, dtype=np.float64), dt=dt) as s:
            s.set_state(qn, pn)
            r = s.lj_trajectory(eps=float(eps), sigma=float(sigma), grad=True, **opts)
        ctx.save_for_backward(
            torch.as_tensor(r["dloss_dq0"], dtype=q0.dtype, device=q0.device),
            torch.tensor(r["dloss_deps"], dtype=torch.float64, device=q0.device),
            torch.tensor(r["dloss_dsigma"], dtype=torch.float64, device=q0.device))
        return torch.tensor(r["loss"], dtype=torch.float64, device=q0.device)

    @staticmethod
    def backward(ctx, g):
        dq, de, ds = ctx.saved_tensors
        return g * dq, (g * de).to(torch.float64), (g * ds).to(torch.float64), None, None, None
