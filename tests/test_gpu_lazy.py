"""GPU: operators traced by ``HIPDGFEMLazyArrayContext``, matched and run on
the HIP kernels, against the oracle; and the runner's reference command line
on the tracing context.
"""

import numpy as np
import pytest
import torch

from actx_dgfem_suite_amd.containers import ConservedEulerField, WaveState, flatten_fields, obj_array
from actx_dgfem_suite_amd.lazy import HIPDGFEMLazyArrayContext
from actx_dgfem_suite_amd.lazy import operators as lo
from actx_dgfem_suite_amd.lazy.grudge_ops import LazyDiscretizationCollection
from actx_dgfem_suite_amd.plan import DGPlan
from tests._helpers import discr_pair, field_errors, oracle_rhs, random_state

pytestmark = pytest.mark.gpu

PARAM = {"wave": 1.0, "euler": 1.4, "maxwell": 0.5}


def _container(equation, leaves):
    if equation == "wave":
        return WaveState(u=leaves[0], v=obj_array(leaves[1:4]))
    if equation == "euler":
        return ConservedEulerField(mass=leaves[0], energy=leaves[1], momentum=obj_array(leaves[2:5]))
    return obj_array(leaves)


@pytest.mark.parametrize("equation", ["wave", "euler", "maxwell"])
@pytest.mark.parametrize("order", [1, 2, 3, 4])
def test_traced_operator_runs_on_the_kernels(equation, order, device):
    """compile -> trace -> normalise -> match -> DGPlan; the result against
    the oracle (1e-12 per field) and against the plan made directly from the
    discretization (the extracted operators differ from its own by rounding)."""
    print(f"[case] lazy {equation} P{order} n=3", flush=True)
    d, od = discr_pair(order, 3)
    actx = HIPDGFEMLazyArrayContext(device)
    dcoll = LazyDiscretizationCollection(actx, d)
    param = PARAM[equation]
    op_fn = {"wave": lo.wave_operator, "euler": lo.euler_operator,
             "maxwell": lo.maxwell_operator}[equation]

    if equation == "wave":
        def rhs(w):
            return op_fn(actx, dcoll, param, w)
    else:
        def rhs(t, w):
            return op_fn(actx, dcoll, param, w)

    compiled = actx.compile(rhs)
    q = random_state(equation, od.x, seed=order)
    state = _container(equation, [torch.from_numpy(np.ascontiguousarray(f)).to(device) for f in q])
    args = (state,) if equation == "wave" else (np.float64(0.5), state)
    out = compiled(*args)
    torch.cuda.synchronize()
    assert compiled.matched.equation == equation
    assert type(out) is type(state)
    got = np.stack([t.cpu().numpy() for t in flatten_fields(out)])
    errs = field_errors(got, oracle_rhs(equation, od, q, param))
    print(f"  rel L2 per field vs oracle: {['%.2e' % e for e in errs]}", flush=True)
    assert max(errs) < 1e-12, errs
    direct = DGPlan(d, equation, param, device=device)
    ref = np.stack([t.cpu().numpy() for t in direct.rhs(flatten_fields(state))])
    assert max(field_errors(got, ref)) < 1e-13
    # a second call reuses the plan
    out2 = compiled(*args)
    torch.cuda.synchronize()
    assert np.array_equal(np.stack([t.cpu().numpy() for t in flatten_fields(out2)]), got)


def test_runner_on_the_tracing_context(device):
    """The reference's command line with --actxs hip:dgfem_lazy: every
    tiny_* equation at P1-P4 in 3D, with the value check against the
    LDS-staged kernel."""
    print("[case] runner hip:dgfem_lazy tiny_* 3D P1-P4", flush=True)
    from actx_dgfem_suite_amd.run.__main__ import _NAME_TO_ACTX_CLASS, main
    rates = main(equations=["tiny_wave", "tiny_euler", "tiny_maxwell"], dims=[3],
                 degrees=[1, 2, 3, 4], actx_ts=[_NAME_TO_ACTX_CLASS["hip:dgfem_lazy"]],
                 verify=True)
    assert np.all(np.isfinite(rates)) and np.all(rates > 0)
