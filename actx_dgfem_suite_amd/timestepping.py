"""Low-storage RK4 time stepping on the device, forward and reverse.

The reference evaluates single right-hand sides only; the LSERK4 scheme is
the one libparanumal's baseline uses (assets/h200nvl.patch:42).  Every update
runs in ``libdgfem.so`` (``dg_lserk_stage`` / ``dg_lserk_stage_adjoint``);
torch supplies device memory and the stream.

Reverse sweep of one step.  With stage inputs q_0 = q^n, q_k and the
recurrences ``res_k = a_k res_{k-1} + dt L(q_{k-1})``,
``q_k = q_{k-1} + b_k res_k`` (k = 1..5), the cotangent of q^n given that of
q^{n+1} follows by transposing each stage in reverse order::

    s = b_5 q_bar
    for k = 5..1:   w = (dL/dq (q_{k-1}))^T s
                    q_bar += dt w;  s = a_k s + b_{k-1} q_bar

(``a_1 = 0`` ends the chain; ``b_0`` is unused).
"""

from __future__ import annotations

from collections.abc import Sequence

import torch

from actx_dgfem_suite_amd.discretization import Discretization

# Carpenter & Kennedy (1994), NASA TM-109112: five-stage, fourth-order,
# 2N-storage Runge-Kutta coefficients.
LSERK4_A = (0.0,
            -567301805773.0 / 1357537059087.0,
            -2404267990393.0 / 2016746695238.0,
            -3550918686646.0 / 2091501179385.0,
            -1275806237668.0 / 842570457699.0)
LSERK4_B = (1432997174477.0 / 9575080441755.0,
            5161836677717.0 / 13612068292357.0,
            1720146321549.0 / 2090206949498.0,
            3134564353537.0 / 4481467310338.0,
            2277821191437.0 / 14882151754819.0)
LSERK4_C = (0.0,
            1432997174477.0 / 9575080441755.0,
            2526269341429.0 / 6820363183857.0,
            2006345519317.0 / 3224310063776.0,
            2802321613138.0 / 2924317926251.0)


def cfl_timestep(discr: Discretization, speed: float, cfl: float = 1.0) -> float:
    """Explicit step ``cfl / (speed * (N+1)^2 * max Fscale)``: the nodal-DG
    estimate with the largest surface-to-volume ratio ``Fscale = sJ/J`` of the
    mesh standing in for 1/h."""
    fmax = float(discr.geometry.fscale.max())
    return cfl / (speed * (discr.order + 1) ** 2 * fmax)


def _clone(fields: Sequence[torch.Tensor]) -> list[torch.Tensor]:
    return [t.clone() for t in fields]


class LSERK4:
    """Forward and reverse LSERK4 steps of one plan, FP64, with buffers
    allocated once."""

    def __init__(self, plan, dt: float):
        self.plan = plan
        self.dt = float(dt)
        self._res = plan.empty_fields()
        self._rhs = plan.empty_fields()
        self._s = plan.empty_fields()
        self._w = plan.empty_fields()

    def step(self, q: Sequence[torch.Tensor],
             stages: list[list[torch.Tensor]] | None = None) -> None:
        """Advance ``q`` in place by one step.  With ``stages`` given, the
        five stage inputs are appended to it (copies)."""
        plan = self.plan
        for a, b in zip(LSERK4_A, LSERK4_B):
            if stages is not None:
                stages.append(_clone(q))
            plan.rhs(q, self._rhs)
            plan.lserk_stage(a, b, self.dt, q, self._res, self._rhs)

    def run(self, q: Sequence[torch.Tensor], nsteps: int) -> None:
        for _ in range(nsteps):
            self.step(q)

    def step_vjp(self, q_n: Sequence[torch.Tensor], q_bar: Sequence[torch.Tensor]) -> None:
        """Replace ``q_bar`` (the cotangent of q^{n+1}) in place by that of
        q^n; ``q_n`` is the state at the start of the step (left intact)."""
        plan = self.plan
        nonlinear = plan.equation == "euler"
        stages: list[list[torch.Tensor]] | None = None
        if nonlinear:
            stages = []
            self.step(_clone(q_n), stages)
        plan.lserk_stage_adjoint(0.0, LSERK4_B[4], self.dt, q_bar, self._s, None)
        for k in range(4, -1, -1):
            plan.rhs_vjp(stages[k] if nonlinear else None, self._s, self._w)
            a = LSERK4_A[k]
            b = LSERK4_B[k - 1] if k > 0 else 0.0
            plan.lserk_stage_adjoint(a, b, self.dt, q_bar, self._s, self._w)


def forward_with_checkpoints(stepper: LSERK4, q0: Sequence[torch.Tensor], nsteps: int,
                             checkpoint_every: int = 10
                             ) -> tuple[list[torch.Tensor], dict[int, list[torch.Tensor]]]:
    """Run ``nsteps`` steps from ``q0`` (left intact), keeping the state at
    step 0 and every ``checkpoint_every`` steps.  Returns ``(q_final,
    checkpoints)``."""
    if checkpoint_every < 1:
        raise ValueError("checkpoint_every must be >= 1")
    q = _clone(q0)
    checkpoints = {0: _clone(q0)}
    for n in range(nsteps):
        stepper.step(q)
        if (n + 1) % checkpoint_every == 0 and n + 1 < nsteps:
            checkpoints[n + 1] = _clone(q)
    return q, checkpoints


def reverse_sweep(stepper: LSERK4, checkpoints: dict[int, list[torch.Tensor]], nsteps: int,
                  q_bar_final: Sequence[torch.Tensor],
                  checkpoint_every: int = 10) -> list[torch.Tensor]:
    """Cotangent of the initial state from that of the final one.  Each
    segment is recomputed from its checkpoint (its step-start states held on
    the device) and its steps are transposed last to first.  Consumes
    ``checkpoints``."""
    lam = _clone(q_bar_final)
    for start in reversed(range(0, nsteps, checkpoint_every)):
        end = min(start + checkpoint_every, nsteps)
        states = [checkpoints.pop(start)]
        q = _clone(states[0])
        for _ in range(start, end - 1):
            stepper.step(q)
            states.append(_clone(q))
        for n in range(end - 1, start - 1, -1):
            stepper.step_vjp(states[n - start], lam)
    return lam


def trajectory_vjp(stepper: LSERK4, q0: Sequence[torch.Tensor], nsteps: int,
                   q_bar_final: Sequence[torch.Tensor],
                   checkpoint_every: int = 10) -> tuple[list[torch.Tensor], list[torch.Tensor]]:
    """``(q_final, q0_bar)`` of an ``nsteps`` trajectory, checkpointed every
    ``checkpoint_every`` steps.  Device memory: ``nsteps / checkpoint_every +
    checkpoint_every`` states."""
    q_final, checkpoints = forward_with_checkpoints(stepper, q0, nsteps, checkpoint_every)
    return q_final, reverse_sweep(stepper, checkpoints, nsteps, q_bar_final, checkpoint_every)
