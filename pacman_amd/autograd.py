"""torch.autograd.Function over the RBF-PUM transfer of a plan.

The transfer is linear in the source values, so its backward is the
transpose apply of the same plan (pacman_rbf_plan_apply_transpose): the
gradient with respect to the source values. The geometry (point
coordinates) is fixed by the plan and receives no gradient.

    plan = pacman_amd.rbf.Plan(3, 0x11, src, tgt, 100, 0.15, -1.0)
    u = pacman_amd.autograd.transfer(plan, values)   # values.requires_grad
    loss = 0.5 * ((u - g) ** 2).sum(); loss.backward()
"""

from __future__ import annotations

import torch


class RbfTransfer(torch.autograd.Function):
    """u = Op(values) with Op the plan's transfer; d loss / d values =
    Op^T (d loss / d u)."""

    @staticmethod
    def forward(ctx, values, plan):
        ctx.plan = plan
        return plan.apply(values.detach().contiguous())

    @staticmethod
    def backward(ctx, grad_out):
        return ctx.plan.apply_transpose(grad_out.contiguous()), None


def transfer(plan, values):
    """Differentiable transfer of values (N,) or (F, N) through plan."""
    return RbfTransfer.apply(values, plan)
