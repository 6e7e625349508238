"""``HIPDGFEMArrayContext``: the array context that runs DG right-hand sides on
gfx950 through ``libdgfem.so``.

It plays the role of the reference's ``DGFEMOptimizerArrayContext``
(arraycontext/dgfem_optimizer.py:63-245) at the plug-in point
``_NAME_TO_ACTX_CLASS`` (run/__main__.py:209-216): ``actx.compile(rhs)``
returns a callable, and ``compiled(*args)`` evaluates the operator.

Matching.  The reference traces ``rhs`` into a pytato DAG and rewrites it to
five canonical batched einsums (arraycontext/materialization_policy.py:
281-293).  Here the operator builders (``actx_dgfem_suite_amd.equations``)
attach a :class:`DGOperator` description to the callable -- equation,
discretization and scalar parameter -- and ``compile`` dispatches on it.  A
callable that carries no such description is not a recognised DG operator
and ``compile`` raises ``NotImplementedError``, the reference's convention for
unmatched DAGs (dgfem_optimizer.py:220-224).

Leaves are ``torch.float64`` tensors of shape ``(K, Np)`` on the context's
device.  Launches go to torch's current stream; ``queue.finish()``
synchronises, mirroring ``finish_command_queue`` (measure.py:48-55).
"""

from __future__ import annotations

from dataclasses import dataclass, field
from types import SimpleNamespace
from typing import Any, Callable

import numpy as np
import torch

from actx_dgfem_suite_amd.containers import flatten_fields, map_leaves, unflatten_like
from actx_dgfem_suite_amd.discretization import Discretization


@dataclass(frozen=True, eq=False)
class DGOperator:
    """The operator an ``rhs`` callable evaluates.

    ``takes_time``: the callable's signature is ``rhs(t, state)`` (Euler,
    Maxwell: equations/euler.py:176, equations/maxwell.py:69-70) rather than
    ``rhs(state)`` (wave: equations/wave.py:281).  The time argument is
    unused by all three operators, as in the reference.
    """
    equation: str
    discr: Discretization
    param: float
    takes_time: bool
    extra: dict = field(default_factory=dict)


class _Queue:
    def __init__(self, device):
        self.device = device

    def finish(self) -> None:
        torch.cuda.synchronize(self.device)


class HIPDGFEMArrayContext:
    """Array context backed by hand-written gfx950 kernels.

    ``precision``: ``"fp64"`` (the drop-in form), or ``"fp32"`` / ``"bf16x3"``
    (the mixed form: state leaves are float32 tensors, ``from_numpy``
    converts).  ``kernel_variant`` picks the FP64 kernel (``"auto"``,
    ``"mfma"``, ``"lds"``).  ``verify_rtol``: the tolerance the runner's value
    check applies to this context.
    """

    kernel_variant = "auto"
    default_precision = "fp64"
    verify_rtol = 1e-5

    def __init__(self, device: torch.device | int | None = None,
                 precision: str | None = None):
        if not torch.cuda.is_available():
            raise RuntimeError("HIPDGFEMArrayContext needs a ROCm GPU (no CPU fallback)")
        self.device = torch.device("cuda", torch.cuda.current_device() if device is None
                                   else torch.device(device).index or 0)
        self.precision = self.default_precision if precision is None else precision
        self.queue = _Queue(self.device)
        self._plans: dict[int, Any] = {}
        self.np = SimpleNamespace(
            exp=torch.exp, sqrt=torch.sqrt, sin=torch.sin, cos=torch.cos,
            abs=torch.abs, maximum=torch.maximum, minimum=torch.minimum,
            real=lambda x: map_leaves(lambda t: t.real if t.is_complex() else t, x),
            where=torch.where, zeros_like=torch.zeros_like, ones_like=torch.ones_like,
        )

    # {{{ data movement

    @property
    def state_dtype(self) -> torch.dtype:
        return torch.float64 if self.precision == "fp64" else torch.float32

    def from_numpy(self, x):
        dtype = self.state_dtype

        def leaf(a):
            t = torch.as_tensor(np.ascontiguousarray(a), device=self.device)
            return t.to(dtype) if t.is_floating_point() else t

        return map_leaves(leaf, x)

    def to_numpy(self, x):
        return map_leaves(lambda t: t.detach().cpu().numpy(), x)

    def freeze(self, x):
        return x

    def thaw(self, x):
        return x

    def freeze_thaw(self, x):
        return x

    def zeros(self, shape, dtype=torch.float64):
        return torch.zeros(shape, dtype=dtype, device=self.device)

    def empty(self, shape, dtype=torch.float64):
        return torch.empty(shape, dtype=dtype, device=self.device)

    # }}}

    def plan_for(self, op: DGOperator):
        key = id(op)
        entry = self._plans.get(key)
        if entry is None:
            from actx_dgfem_suite_amd.plan import DGPlan
            entry = (op, DGPlan(op.discr, op.equation, op.param, device=self.device,
                                variant=self.kernel_variant))
            self._plans[key] = entry
        return entry[1]

    def apply_operator(self, op: DGOperator, state):
        plan = self.plan_for(op)
        leaves = flatten_fields(state)
        out = plan.rhs(leaves, precision=self.precision)
        return unflatten_like(state, out)

    def compile(self, f: Callable) -> Callable:
        op = getattr(f, "dg_operator", None)
        if not isinstance(op, DGOperator):
            raise NotImplementedError(
                f"{f!r} does not describe a DG operator this context can map to "
                "its kernels (wave/Euler/Maxwell built by "
                "actx_dgfem_suite_amd.equations)")
        self.plan_for(op)

        def compiled(*args):
            if op.takes_time:
                if len(args) != 2:
                    raise TypeError("expected (t, state)")
                state = args[1]
            else:
                if len(args) != 1:
                    raise TypeError("expected (state,)")
                state = args[0]
            return self.apply_operator(op, state)

        compiled.dg_operator = op
        return compiled

    def clone(self) -> "HIPDGFEMArrayContext":
        return type(self)(device=self.device, precision=self.precision)


class HIPDGFEMLDSArrayContext(HIPDGFEMArrayContext):
    """FP64 drop-in form on the LDS-staged baseline kernel."""
    kernel_variant = "lds"


class HIPDGFEMFP32ArrayContext(HIPDGFEMArrayContext):
    """Mixed form: FP32 state, FP32 MFMA contractions, FP64 geometry and flux.
    Its value check allows the FP32 state's rounding (relative 1e-3)."""
    default_precision = "fp32"
    verify_rtol = 1e-3


class HIPDGFEMBF16x3ArrayContext(HIPDGFEMFP32ArrayContext):
    """Mixed form with the contractions as a three-product bf16 split."""
    default_precision = "bf16x3"
