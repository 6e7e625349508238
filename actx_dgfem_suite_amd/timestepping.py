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


# Forms whose fused step measured faster on MI355X (20 whole steps through
# ``run`` on the 4M-DOF meshes, 94k-1.05M tets, fused against the RHS and
# stage kernels).  FP64, at least 1%: wave P1 1.25x, P3 1.19x, Maxwell P1
# 1.24x, P2 1.01x, Euler P2 1.02x; wave P2 1.00x, P4 0.98x, Euler P1 0.93x,
# P3 0.95x, P4 1.01x, Maxwell P3 1.00x, P4 1.00x.  Mixed, at least 4%: FP32
# wave P1 1.15x, Euler P1 1.08x, P2 1.11x, Maxwell P2 1.04x, P3 1.08x (Euler
# P3 1.03x, P4 0.96x, Maxwell P1 0.99x, P4 0.97x); BF16x3 wave P1 1.06x, P2
# 1.04x, P4 1.06x on 584k tets (1.00x on 118k), Euler P3 1.08x (Maxwell P3
# 0.87x); elsewhere 0.63x-1.02x.
# Unfused, the stage kernel re-reads the state and RHS that the RHS kernel
# has just touched while they are still in the 256 MB Infinity Cache, and the
# fused forms' stage epilogue costs registers their RHS kernels are short of.
_FUSED_FP64 = {("wave", 1), ("wave", 3), ("euler", 2), ("maxwell", 1), ("maxwell", 2)}
_FUSED_MIXED = {
    "fp32": {("wave", 1), ("euler", 1), ("euler", 2), ("maxwell", 2), ("maxwell", 3)},
    "bf16x3": {("wave", 1), ("wave", 2), ("wave", 4), ("euler", 3)},
}


# Linear operators whose fused reverse step in the adjoint variables
# (``dg_rhs_lserk_stage_adjoint``) measured at least 3% faster on MI355X
# than ``dg_rhs`` + ``dg_lserk_stage_adjoint`` (20 reverse steps on the
# 4M-DOF meshes, 94k-1.05M tets): wave P1 1.24x, P3 1.16x, Maxwell P1
# 1.23x, P2 1.05x.  Wave P2 1.01x, P4 0.99x, Maxwell P3 0.98x, P4 1.00x.
_FUSED_REVERSE = {("wave", 1), ("wave", 3), ("maxwell", 1), ("maxwell", 2)}


def fused_stage_preferred(plan, precision: str = "fp64") -> bool:
    """Default of ``LSERK4(fused=None)``: fused stages on the tetrahedral
    cases of ``_FUSED_FP64`` (FP64) and ``_FUSED_MIXED`` (mixed forms), where
    the plan has them."""
    if precision != "fp64":
        return bool(plan.fused_stage_mixed) and \
            (plan.equation, plan.discr.order) in _FUSED_MIXED[precision]
    return bool(plan.fused_stage) and (plan.equation, plan.discr.order) in _FUSED_FP64


class LSERK4:
    """Forward and reverse LSERK4 steps of one plan, with buffers allocated
    once.

    ``precision`` is the forward RHS: ``"fp64"``, or the mixed ``"fp32"`` /
    ``"bf16x3"`` forms, which keep the state and residual in FP64 and feed the
    RHS an FP32 shadow of the state (``dg_lserk_stage_mixed``).  The reverse
    step runs in the dtype of the cotangent it is given: float64 or float32.
    The linear operators' reverse steps run in the variables of
    ``plan.adjoint_space`` on the forward RHS (FP64, or the FP32 mixed
    kernel); Euler's on the VJP (``dg_rhs_vjp``, or the FP32 adjoint
    ``dg_rhs_vjp_fp32``).
    """

    def __init__(self, plan, dt: float, precision: str = "fp64", fused: bool | None = None):
        if precision not in ("fp64", "fp32", "bf16x3"):
            raise ValueError(f"precision '{precision}'")
        available = plan.fused_stage if precision == "fp64" else plan.fused_stage_mixed
        if fused and not available:
            raise ValueError("fused stages need plan.fused_stage (FP64) or "
                             "plan.fused_stage_mixed (mixed forms)")
        self.plan = plan
        self.dt = float(dt)
        self.precision = precision
        self._res = plan.empty_fields()
        explicit = fused
        if fused is None:
            fused = fused_stage_preferred(plan, precision)
        # fused: the stage runs inside the RHS kernel, which reads one state
        # buffer (FP64) or FP32 shadow (mixed) and writes the other
        self._q2 = plan.empty_fields() if fused and precision == "fp64" else None
        self._q32b = plan.empty_fields(torch.float32) if fused and precision != "fp64" else None
        # float64 reverse steps of the linear operators in the adjoint
        # variables: fused on an FP64 stepper's explicit choice, else where
        # it measured faster
        fwd64 = explicit if precision == "fp64" else None
        self._fused_adjoint = (plan.equation != "euler" and bool(plan.fused_stage) and
                               (bool(fwd64) if fwd64 is not None else
                                (plan.equation, plan.discr.order) in _FUSED_REVERSE))
        if precision == "fp64":
            self._rhs = None if self._q2 is not None else plan.empty_fields()
            self._q32 = None
        else:
            self._rhs = None if self._q32b is not None else plan.empty_fields(torch.float32)
            self._q32 = plan.empty_fields(torch.float32)
        self._adj: dict[torch.dtype, tuple[list[torch.Tensor], list[torch.Tensor]]] = {}

    @property
    def linear(self) -> bool:
        """The operator is linear: its VJP does not read the state."""
        return self.plan.equation != "euler"

    def step(self, q: Sequence[torch.Tensor],
             stages: list[list[torch.Tensor]] | None = None) -> None:
        """Advance the FP64 state ``q`` in place by one step.  With ``stages``
        given, the five stage inputs are appended to it (copies)."""
        plan = self.plan
        if self._q2 is not None:
            self._fused_steps(q, 1, stages)
            return
        if self._q32 is None:
            for a, b in zip(LSERK4_A, LSERK4_B):
                if stages is not None:
                    stages.append(_clone(q))
                plan.rhs(q, self._rhs)
                plan.lserk_stage(a, b, self.dt, q, self._res, self._rhs)
            return
        for t, src in zip(self._q32, q):
            t.copy_(src)
        if self._q32b is not None:
            # the FP64 state and residual are updated in place; the kernel
            # gathers one FP32 shadow and writes the other
            s32, d32 = self._q32, self._q32b
            for a, b in zip(LSERK4_A, LSERK4_B):
                if stages is not None:
                    stages.append(_clone(q))
                plan.rhs_lserk_stage_mixed(self.precision, a, b, self.dt, s32, d32, q, self._res)
                s32, d32 = d32, s32
            return
        for a, b in zip(LSERK4_A, LSERK4_B):
            if stages is not None:
                stages.append(_clone(q))
            plan.rhs_mixed(self._q32, self._rhs, self.precision)
            plan.lserk_stage_mixed(a, b, self.dt, q, self._res, self._rhs, self._q32)

    @property
    def fused(self) -> bool:
        """Whether the steps run fused stages (``dg_rhs_lserk_stage`` or
        ``dg_rhs_lserk_stage_mixed``)."""
        return self._q2 is not None or self._q32b is not None

    def _fused_steps(self, q: Sequence[torch.Tensor], nsteps: int,
                     stages: list[list[torch.Tensor]] | None = None) -> None:
        # dg_rhs_lserk_stage alternates q and the second state buffer; after
        # an odd number of stages the state is in the second one and is
        # copied back
        src, dst = list(q), self._q2
        for _ in range(nsteps):
            for a, b in zip(LSERK4_A, LSERK4_B):
                if stages is not None:
                    stages.append(_clone(src))
                self.plan.rhs_lserk_stage(a, b, self.dt, src, dst, self._res)
                src, dst = dst, src
        if src is self._q2:
            for t, s in zip(q, src):
                t.copy_(s)

    def run(self, q: Sequence[torch.Tensor], nsteps: int) -> None:
        """Advance ``q`` in place by ``nsteps`` steps (fused: one copy at the
        end when the stage count is odd, not one per step)."""
        if self._q2 is not None:
            self._fused_steps(q, nsteps)
            return
        for _ in range(nsteps):
            self.step(q)

    def _adjoint_buffers(self, dtype):
        buf = self._adj.get(dtype)
        if buf is None:
            buf = self._adj[dtype] = (self.plan.empty_fields(dtype), self.plan.empty_fields(dtype))
        return buf

    def _adjoint_space_sweep(self, q_hat: Sequence[torch.Tensor]) -> bool:
        """Whether a cotangent of this stepper is swept in the variables of
        ``plan.adjoint_space``: linear operators, float64 (on the FP64 RHS)
        or float32 (on the FP32 mixed RHS)."""
        return self.linear and q_hat[0].dtype in (torch.float64, torch.float32)

    def step_vjp_adjoint_space(self, q_hat: Sequence[torch.Tensor]) -> None:
        """One reverse step of a linear operator on a cotangent held in the
        variables ``q_hat = R M^{-1} q_bar`` of ``plan.adjoint_space``.  There
        the transposed RHS is the forward RHS (``L^T = R M L R M^{-1}``), so
        each reverse stage is one forward RHS and one reverse stage update
        (float64: ``dg_rhs``, ``dg_lserk_stage_adjoint``; float32: the FP32
        mixed ``dg_rhs_mixed``, ``dg_lserk_stage_adjoint_fp32``): the reverse
        step costs what a forward step does."""
        plan = self.plan
        dtype = q_hat[0].dtype
        s, w = self._adjoint_buffers(dtype)
        precision = "fp64" if dtype == torch.float64 else "fp32"
        plan.lserk_stage_adjoint(0.0, LSERK4_B[4], self.dt, q_hat, s, None)
        if dtype == torch.float64 and self._fused_adjoint:
            # dg_rhs_lserk_stage_adjoint: q_hat in place, s alternating
            # between the two scratch buffers
            src, dst = s, w
            for k in range(4, -1, -1):
                b = LSERK4_B[k - 1] if k > 0 else 0.0
                plan.rhs_lserk_stage_adjoint(LSERK4_A[k], b, self.dt, src, dst, q_hat)
                src, dst = dst, src
            return
        for k in range(4, -1, -1):
            plan.rhs(s, w, precision=precision)
            a = LSERK4_A[k]
            b = LSERK4_B[k - 1] if k > 0 else 0.0
            plan.lserk_stage_adjoint(a, b, self.dt, q_hat, s, w)

    def step_vjp(self, q_n: Sequence[torch.Tensor] | None, q_bar: Sequence[torch.Tensor]) -> None:
        """Replace ``q_bar`` (the cotangent of q^{n+1}) in place by that of
        q^n.  ``q_n`` is the FP64 state at the start of the step (left
        intact); linear operators do not read it and accept ``None``."""
        plan = self.plan
        dtype = q_bar[0].dtype
        if self._adjoint_space_sweep(q_bar):
            plan.adjoint_space(q_bar, True, out=q_bar)
            self.step_vjp_adjoint_space(q_bar)
            plan.adjoint_space(q_bar, False, out=q_bar)
            return
        s, w = self._adjoint_buffers(dtype)
        stages: list[list[torch.Tensor]] | None = None
        if not self.linear:
            if q_n is None:
                raise ValueError("step_vjp: the Euler adjoint needs the step's start state")
            stages = []
            self.step(_clone(q_n), stages)
            if dtype != torch.float64:
                stages = [[t.to(dtype) for t in st] for st in stages]
        plan.lserk_stage_adjoint(0.0, LSERK4_B[4], self.dt, q_bar, s, None)
        for k in range(4, -1, -1):
            plan.rhs_vjp(None if stages is None else stages[k], s, w)
            a = LSERK4_A[k]
            b = LSERK4_B[k - 1] if k > 0 else 0.0
            plan.lserk_stage_adjoint(a, b, self.dt, q_bar, s, w)


def forward_with_checkpoints(stepper: LSERK4, q0: Sequence[torch.Tensor], nsteps: int,
                             checkpoint_every: int = 10
                             ) -> tuple[list[torch.Tensor], dict[int, list[torch.Tensor]]]:
    """Run ``nsteps`` steps from ``q0`` (left intact), keeping the state at
    step 0 and every ``checkpoint_every`` steps -- for the nonlinear Euler
    operator only: the linear operators' adjoint reads no forward state, so
    they keep none.  Returns ``(q_final, checkpoints)``."""
    if checkpoint_every < 1:
        raise ValueError("checkpoint_every must be >= 1")
    keep = not stepper.linear
    q = _clone(q0)
    if not keep:
        stepper.run(q, nsteps)
        return q, {}
    checkpoints = {0: _clone(q0)}
    for start in range(0, nsteps, checkpoint_every):
        end = min(start + checkpoint_every, nsteps)
        stepper.run(q, end - start)
        if end < nsteps:
            checkpoints[end] = _clone(q)
    return q, checkpoints


def reverse_sweep(stepper: LSERK4, checkpoints: dict[int, list[torch.Tensor]], nsteps: int,
                  q_bar_final: Sequence[torch.Tensor],
                  checkpoint_every: int = 10) -> list[torch.Tensor]:
    """Cotangent of the initial state from that of the final one (in the
    final cotangent's dtype).  For the nonlinear Euler operator each segment
    is recomputed from its checkpoint (its step-start states held on the
    device) and its steps are transposed last to first; the linear operators'
    adjoint reads no state, so their steps are transposed directly.  Consumes
    ``checkpoints``."""
    lam = _clone(q_bar_final)
    if stepper.linear:
        checkpoints.clear()
        if stepper._adjoint_space_sweep(lam):
            # the whole sweep in the adjoint variables: one transform each way
            stepper.plan.adjoint_space(lam, True, out=lam)
            for _ in range(nsteps):
                stepper.step_vjp_adjoint_space(lam)
            stepper.plan.adjoint_space(lam, False, out=lam)
            return lam
        for _ in range(nsteps):
            stepper.step_vjp(None, lam)
        return lam
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


def misfit_gradient(stepper: LSERK4, q0: Sequence[torch.Tensor], target: Sequence[torch.Tensor],
                    nsteps: int, checkpoint_every: int = 10,
                    adjoint_dtype: torch.dtype = torch.float64
                    ) -> tuple[float, list[torch.Tensor]]:
    """``J = 1/2 sum_c ||q_c(T) - target_c||^2`` over the first
    ``len(target)`` fields after ``nsteps`` steps from ``q0``, and its
    gradient with respect to every field of ``q0`` (in ``adjoint_dtype``).
    The forward runs in the stepper's precision."""
    q_final, checkpoints = forward_with_checkpoints(stepper, q0, nsteps, checkpoint_every)
    diff = [a - b for a, b in zip(q_final, target)]
    j = 0.5 * float(sum(torch.sum(d * d) for d in diff))
    q_bar = [(diff[c] if c < len(diff) else torch.zeros_like(q_final[c])).to(adjoint_dtype)
             for c in range(len(q_final))]
    return j, reverse_sweep(stepper, checkpoints, nsteps, q_bar, checkpoint_every)
