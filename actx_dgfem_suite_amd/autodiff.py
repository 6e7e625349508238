"""torch.autograd bindings of the adjoint kernels.

``rhs(plan, q)`` is one right-hand side evaluation whose backward is
``dg_rhs_vjp``; ``trajectory(stepper, q0, nsteps, checkpoint_every)`` is an
LSERK4 trajectory whose backward is the checkpointed reverse sweep of
:mod:`actx_dgfem_suite_amd.timestepping`.  Both run FP64 fields on the plan's
device.  The reference's counterpart would be ``jax.vjp`` of the RHS compiled
by its jax:jit array context (run/__main__.py:213).
"""

from __future__ import annotations

from collections.abc import Sequence

import torch

from actx_dgfem_suite_amd.timestepping import LSERK4, forward_with_checkpoints, reverse_sweep


def _cotangents(plan, grads) -> list[torch.Tensor]:
    return [torch.zeros(plan.shape, dtype=torch.float64, device=plan.device) if g is None
            else g.contiguous() for g in grads]


class _RHS(torch.autograd.Function):
    @staticmethod
    def forward(ctx, plan, *q):
        ctx.plan = plan
        if plan.equation == "euler":
            ctx.save_for_backward(*q)
        return tuple(plan.rhs([t.contiguous() for t in q]))

    @staticmethod
    def backward(ctx, *grads):
        plan = ctx.plan
        q = list(ctx.saved_tensors) if plan.equation == "euler" else None
        return (None, *plan.rhs_vjp(q, _cotangents(plan, grads)))


class _Trajectory(torch.autograd.Function):
    @staticmethod
    def forward(ctx, stepper, nsteps, checkpoint_every, *q0):
        q0 = [t.detach().contiguous() for t in q0]
        q_final, checkpoints = forward_with_checkpoints(stepper, q0, nsteps, checkpoint_every)
        ctx.stepper = stepper
        ctx.nsteps = nsteps
        ctx.checkpoint_every = checkpoint_every
        ctx.checkpoints = checkpoints
        return tuple(q_final)

    @staticmethod
    def backward(ctx, *grads):
        if ctx.checkpoints is None:
            raise RuntimeError("trajectory backward called twice (checkpoints consumed)")
        plan = ctx.stepper.plan
        lam = reverse_sweep(ctx.stepper, ctx.checkpoints, ctx.nsteps,
                            _cotangents(plan, grads), ctx.checkpoint_every)
        ctx.checkpoints = None
        return (None, None, None, *lam)


def rhs(plan, q: Sequence[torch.Tensor]) -> list[torch.Tensor]:
    """Differentiable ``L(q)`` (FP64)."""
    return list(_RHS.apply(plan, *q))


def trajectory(stepper: LSERK4, q0: Sequence[torch.Tensor], nsteps: int,
               checkpoint_every: int = 10) -> list[torch.Tensor]:
    """Differentiable state after ``nsteps`` LSERK4 steps from ``q0``."""
    return list(_Trajectory.apply(stepper, int(nsteps), int(checkpoint_every), *q0))
