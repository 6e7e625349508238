"""The array-context boundary: actx.compile(rhs)(*args) as the reference's
runner calls it (run/__main__.py:72-76), for all three operator builders."""

import numpy as np
import pytest

from actx_dgfem_suite_amd.containers import ConservedEulerField, WaveState, flatten_fields
from actx_dgfem_suite_amd.rhs_builder import get_rhs
from oracle import dg_oracle
from tests._helpers import field_errors, oracle_rhs

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("equation", ["tiny_wave", "tiny_euler", "tiny_maxwell"])
@pytest.mark.parametrize("degree", [1, 3])
def test_compiled_rhs_matches_oracle(equation, degree, device):
    from actx_dgfem_suite_amd.arraycontext import HIPDGFEMArrayContext
    print(f"[case] actx {equation} P{degree}", flush=True)
    actx = HIPDGFEMArrayContext()
    rhs, args = get_rhs(equation, actx, 3, degree)
    compiled = actx.compile(rhs)
    out = compiled(*args)
    actx.queue.finish()
    state = args[-1]
    assert type(out) is type(state)
    if equation == "tiny_wave":
        assert isinstance(out, WaveState)
    if equation == "tiny_euler":
        assert isinstance(out, ConservedEulerField)
    op = rhs.dg_operator
    d = op.discr
    od = dg_oracle.build_discretization(d.mesh.vertices, d.mesh.elements, d.ref.nodes)
    q = np.stack(flatten_fields(actx.to_numpy(state)))
    ref = oracle_rhs(op.equation, od, q, op.param)
    got = np.stack(flatten_fields(actx.to_numpy(out)))
    errs = field_errors(got, ref)
    assert max(errs) < 1e-12, errs
    # eager call of the uncompiled operator gives the same result
    eager = np.stack(flatten_fields(actx.to_numpy(rhs(*args))))
    np.testing.assert_array_equal(eager, got)


def test_compile_rejects_unknown_callables(device):
    from actx_dgfem_suite_amd.arraycontext import HIPDGFEMArrayContext
    actx = HIPDGFEMArrayContext()
    with pytest.raises(NotImplementedError):
        actx.compile(lambda w: w)


@pytest.mark.parametrize("name", ["hip:dgfem", "hip:dgfem_lds", "hip:dgfem_fp32",
                                  "hip:dgfem_bf16x3"])
def test_runner_times_every_context(name, device, capsys):
    """The benchmark runner (the reference's run/__main__.py protocol) on the
    tiny tiers, with its value check on."""
    from actx_dgfem_suite_amd.run.__main__ import _NAME_TO_ACTX_CLASS, main
    print(f"[case] runner {name}", flush=True)
    rates = main(equations=["tiny_wave", "tiny_euler", "tiny_maxwell"], dims=[3],
                 degrees=[1, 2], actx_ts=[_NAME_TO_ACTX_CLASS[name]], verify=True)
    assert np.all(rates > 0)
    out = capsys.readouterr().out
    assert "GFLOPS/s for 3D-tiny_maxwell" in out and name in out


@pytest.mark.parametrize("precision", ["fp32", "bf16x3"])
def test_mixed_context_keeps_float32_state(precision, device):
    import torch

    from actx_dgfem_suite_amd.arraycontext import (
        HIPDGFEMBF16x3ArrayContext,
        HIPDGFEMFP32ArrayContext,
    )
    print(f"[case] mixed actx {precision}", flush=True)
    cls = HIPDGFEMFP32ArrayContext if precision == "fp32" else HIPDGFEMBF16x3ArrayContext
    actx = cls()
    rhs, args = get_rhs("tiny_wave", actx, 3, 2)
    out = actx.compile(rhs)(*args)
    leaves = flatten_fields(out)
    assert all(t.dtype == torch.float32 for t in leaves)
    op = rhs.dg_operator
    d = op.discr
    od = dg_oracle.build_discretization(d.mesh.vertices, d.mesh.elements, d.ref.nodes)
    q = np.stack([np.asarray(a, dtype=np.float64) for a in flatten_fields(actx.to_numpy(args[-1]))])
    ref = oracle_rhs(op.equation, od, q, op.param)
    got = np.stack(flatten_fields(actx.to_numpy(out))).astype(np.float64)
    assert max(field_errors(got, ref)) < 1e-4


def test_dof_scaling_tool_runs(device, capsys):
    from actx_dgfem_suite_amd.tools.wave_dof_scaling import main
    print("[case] dof scaling tool", flush=True)
    t = main(dim=3, degrees=[1, 2], ndofs_list=[1000, 20000])
    assert t.shape == (2, 2) and np.all(t > 0)
    assert "GDOFS/s for 3D-wave" in capsys.readouterr().out
