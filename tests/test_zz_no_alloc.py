"""GPU: dg_rhs_vjp_noalloc, the VJP whose workspace is the library's static
device array, against dg_rhs_vjp with a caller workspace.

Two calls in a row, the second on a smaller plan, on the fused MFMA path
(wave, Maxwell) and on the adjoint-kernel path (Euler); both must equal
dg_rhs_vjp bit for bit, since the kernels and their inputs are the same and
only the workspace address differs.  A plan whose workspace exceeds the
array is refused before anything is launched.
"""

import ctypes

import numpy as np
import pytest
import torch

from actx_dgfem_suite_amd import _lib
from actx_dgfem_suite_amd.plan import DGPlan
from tests._helpers import discr_pair, random_state

pytestmark = pytest.mark.gpu

PARAM = {"wave": 1.0, "euler": 1.4, "maxwell": 0.5}
SCRATCH_BYTES = 128 << 20


def _noalloc(plan, q, y, out, device):
    p = _lib.pointer_array
    stream = ctypes.c_void_p(torch.cuda.current_stream(device).cuda_stream)
    return plan.lib.dg_rhs_vjp_noalloc(
        plan._handle, None if q is None else p([t.data_ptr() for t in q]),
        p([t.data_ptr() for t in y]), p([t.data_ptr() for t in out]), stream)


def _case(equation, order, nel, seed, device):
    d, od = discr_pair(order, nel)
    plan = DGPlan(d, equation, PARAM[equation], device=device)
    rng = np.random.default_rng(seed)
    q = None
    if equation == "euler":
        q = [torch.from_numpy(np.ascontiguousarray(f)).to(device)
             for f in random_state("euler", od.x, seed=seed)]
    y = [torch.from_numpy(f).to(device)
         for f in rng.standard_normal((plan.nfields, d.nelements, d.np))]
    return plan, q, y


@pytest.mark.parametrize("first,second", [
    (("maxwell", 3, 4), ("wave", 2, 2)),
    (("euler", 2, 4), ("euler", 1, 2)),
])
def test_noalloc_vjp_matches_dg_rhs_vjp(first, second, device):
    for k, (equation, order, nel) in enumerate((first, second)):
        print(f"[case] noalloc vjp call {k + 1}: {equation} P{order} n={nel}", flush=True)
        plan, q, y = _case(equation, order, nel, 10 + k, device)
        ref = [t.clone() for t in plan.rhs_vjp(q, y)]
        out = plan.empty_fields()
        torch.cuda.synchronize()
        rc = _noalloc(plan, q, y, out, device)
        assert rc == 0, plan.lib.dg_last_error_string()
        torch.cuda.synchronize()
        for a, b in zip(out, ref):
            assert torch.equal(a, b)
        # in place on the cotangent, as dg_rhs_vjp allows
        y2 = [t.clone() for t in y]
        assert _noalloc(plan, q, y2, y2, device) == 0
        torch.cuda.synchronize()
        for a, b in zip(y2, ref):
            assert torch.equal(a, b)


def test_noalloc_vjp_refuses_oversized_and_unsupported(device):
    """Maxwell P4 on 10^3 x 6 tets needs 6 x 6000 x 35 doubles (10 MB): fits;
    on 25^3 x 6 tets it needs 157 MB, more than the 128 MiB array."""
    print("[case] noalloc vjp size limit", flush=True)
    from actx_dgfem_suite_amd.discretization import make_box_discretization
    d = make_box_discretization(4, 25)
    need = 6 * d.nelements * d.np * 8
    assert need > SCRATCH_BYTES
    plan = DGPlan(d, "maxwell", 0.5, device=device)
    y = plan.empty_fields()
    out = plan.empty_fields()
    assert _noalloc(plan, None, y, out, device) == 1
    msg = plan.lib.dg_last_error_string().decode()
    assert "128 MiB" in msg or str(SCRATCH_BYTES) in msg, msg
    # aliasing is checked as in dg_rhs_vjp
    small, q, ys = _case("euler", 1, 2, 3, device)
    assert _noalloc(small, q, ys, q, device) == 1
    assert "q_bar must not overlap q" in small.lib.dg_last_error_string().decode()
