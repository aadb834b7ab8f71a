"""torch.autograd.Function over the RBF-PUM transfer of a plan.

The transfer is linear in the source values, so its backward with respect
to them is the transpose apply of the same plan
(pacman_rbf_plan_apply_transpose). With respect to the target coordinates
the backward is pacman_rbf_plan_target_coord_grad: the analytic derivative
of u(y) = sum_k w_k(y) t_k(y) (Wendland derivatives of the basis and of the
partition-of-unity weights) at the plan's targets, the patch structure
(centres, radius, membership) held fixed. The source coordinates receive no
gradient.

    plan = pacman_amd.rbf.Plan(3, 0x11, src, tgt, 100, 0.15, -1.0)
    tgt.requires_grad_()                       # the plan's target points
    u = pacman_amd.autograd.transfer(plan, values, tgt)
    loss = 0.5 * ((u - g) ** 2).sum(); loss.backward()
"""

from __future__ import annotations

import torch


class RbfTransfer(torch.autograd.Function):
    """u = Op(values) at the plan's targets. d loss / d values =
    Op^T (d loss / d u); d loss / d targets from the saved per-patch
    coefficients."""

    @staticmethod
    def forward(ctx, values, target_points, plan):
        ctx.plan = plan
        v = values.detach().contiguous()
        if target_points is not None and ctx.needs_input_grad[1]:
            if tuple(target_points.shape) != (plan.num_targets, plan.dim):
                raise ValueError("target_points must be the plan's (M, dim) "
                                 "targets")
            out, state = plan.apply_with_state(v)
            ctx.save_for_backward(state)
        else:
            out = plan.apply(v)
        return out

    @staticmethod
    def backward(ctx, grad_out):
        g = grad_out.contiguous()
        gv = gt = None
        if ctx.needs_input_grad[0]:
            gv = ctx.plan.apply_transpose(g)
        if ctx.needs_input_grad[1]:
            (state,) = ctx.saved_tensors
            gt = ctx.plan.target_coord_grad(state, g)
        return gv, gt, None


def transfer(plan, values, target_points=None):
    """Differentiable transfer of values (N,) or (F, N) through plan.

    target_points, when given, must be the (M, dim) tensor of target
    coordinates the plan was built on; it then receives the gradient with
    respect to the target coordinates.
    """
    return RbfTransfer.apply(values, target_points, plan)
