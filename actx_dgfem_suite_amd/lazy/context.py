"""``HIPDGFEMLazyArrayContext``: an array context that compiles a traced DG
operator to the HIP kernels.

It is the part of the reference's ``DGFEMOptimizerArrayContext``
(arraycontext/dgfem_optimizer.py:63-245) that this path needs:
- ``compile(f)`` returns a callable;
- on its first call, the callable traces ``f`` with placeholders for the
  state's leaves;
- the operator code's ``actx.einsum``, ``actx.np.*`` and arithmetic build
  the DAG of :mod:`.ir`;
- the DAG is normalised as the reference does (:mod:`.normalize`) and matched
  to an equation, a parameter and the plan's arrays (:mod:`.match`);
- every call then runs the kernel of that plan on the call's device tensors.

A callable whose DAG does not match raises ``NotImplementedError`` on its
first call, like an unmatched DAG in the reference
(arraycontext/dgfem_optimizer.py:220-224).  There is no host evaluation of
the DAG: the matched plan is the only way a compiled callable computes.
"""

from __future__ import annotations

from types import SimpleNamespace

import numpy as np
import torch

from actx_dgfem_suite_amd.containers import flatten_fields, map_leaves, unflatten_like
from actx_dgfem_suite_amd.lazy import ir
from actx_dgfem_suite_amd.lazy.match import match_dg_operator
from actx_dgfem_suite_amd.lazy.normalize import normalize


def _is_container(x) -> bool:
    try:
        flatten_fields(x)
    except TypeError:
        return False
    return not isinstance(x, (list, tuple)) or all(isinstance(a, torch.Tensor) for a in x)


class HIPDGFEMLazyArrayContext:
    """Traces, matches and compiles DG operators built on
    :mod:`actx_dgfem_suite_amd.lazy.grudge_ops`; state leaves are float64
    ``(K, Np)`` device tensors."""

    # the runner's value check tolerance (run/__main__.py:77-96)
    verify_rtol = 1e-5

    def __init__(self, device: torch.device | int | None = None):
        self.device = None
        if torch.cuda.is_available():
            self.device = torch.device("cuda", torch.cuda.current_device() if device is None
                                       else torch.device(device).index or 0)
        self.queue = SimpleNamespace(finish=self._finish)
        self.np = SimpleNamespace(
            sqrt=lambda x: ir.elementwise("sqrt", x),
            abs=lambda x: ir.elementwise("abs", x),
            maximum=lambda a, b: ir.elementwise("maximum", a, b),
            minimum=lambda a, b: ir.elementwise("minimum", a, b),
            where=lambda c, a, b: ir.elementwise("where", c, a, b),
        )

    def _finish(self) -> None:
        if self.device is not None:
            torch.cuda.synchronize(self.device)

    # {{{ array-context interface used by the operators and the runner

    def einsum(self, spec: str, *args, arg_names=None, tagged=()):
        return ir.Einsum(spec, tuple(ir.as_array(a) for a in args))

    def from_numpy(self, x):
        if self.device is None:
            raise RuntimeError("HIPDGFEMLazyArrayContext needs a ROCm GPU for device arrays "
                               "(no CPU fallback)")
        return map_leaves(lambda a: torch.as_tensor(np.ascontiguousarray(a),
                                                    device=self.device), x)

    def to_numpy(self, x):
        return map_leaves(lambda t: t.detach().cpu().numpy(), x)

    def freeze(self, x):
        return x

    def thaw(self, x):
        return x

    def freeze_thaw(self, x):
        return x

    def tag_axis(self, axis, tags, x):
        return x

    # }}}

    def trace(self, f, args):
        """Placeholders for the state leaves of ``args``; ``f`` applied to them.
        Returns (state position, placeholders, output container, output leaves)."""
        pos = [i for i, a in enumerate(args) if _is_container(a)]
        if len(pos) != 1:
            raise NotImplementedError("expected exactly one state container among the arguments")
        state = args[pos[0]]
        leaves = flatten_fields(state)
        phs = [ir.Placeholder(f"q{k}", tuple(t.shape)) for k, t in enumerate(leaves)]
        traced = list(args)
        traced[pos[0]] = unflatten_like(state, phs)
        out = f(*traced)
        return pos[0], phs, out, flatten_fields(out)

    def compile(self, f):
        return LazyCompiled(self, f)


class LazyCompiled:
    """The callable ``compile`` returns; traces and matches on first call."""

    def __init__(self, actx: HIPDGFEMLazyArrayContext, f):
        self.actx = actx
        self.f = f
        self.plan = None
        self.matched = None
        self._state_pos = None
        self._out_template = None

    def build(self, args):
        pos, phs, out, out_leaves = self.actx.trace(self.f, args)
        if any(not isinstance(o, ir.Array) for o in out_leaves):
            raise NotImplementedError("the traced function does not return array expressions")
        self.matched = match_dg_operator(normalize(out_leaves), phs)
        from actx_dgfem_suite_amd.plan import DGPlan
        self.plan = DGPlan(self.matched.discretization(), self.matched.equation,
                           self.matched.param, device=self.actx.device)
        self._state_pos = pos
        self._out_template = out

    def __call__(self, *args):
        if self.plan is None:
            self.build(args)
        leaves = flatten_fields(args[self._state_pos])
        return unflatten_like(self._out_template, self.plan.rhs(leaves))
