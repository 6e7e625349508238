"""Host-side choices that need no GPU: the LSERK4 stepper's default kernels."""

from types import SimpleNamespace

import pytest
import torch

from actx_dgfem_suite_amd.timestepping import LSERK4, _FUSED_REVERSE


def _mock_plan(equation, order):
    """The attributes LSERK4's constructor reads, on host tensors."""
    return SimpleNamespace(
        equation=equation, discr=SimpleNamespace(order=order), fused_stage=True,
        fused_stage_mixed=True,
        empty_fields=lambda dtype=torch.float64: [torch.empty(4, dtype=dtype) for _ in range(4)])


@pytest.mark.parametrize("equation", ["wave", "maxwell"])
@pytest.mark.parametrize("order", [1, 2, 3, 4])
def test_fused_reverse_default_follows_the_measured_table(equation, order):
    """With ``fused`` left to its default, the float64 reverse stage is fused
    only where it measured faster; an explicit choice wins."""
    plan = _mock_plan(equation, order)
    assert LSERK4(plan, 0.1)._fused_adjoint == ((equation, order) in _FUSED_REVERSE)
    assert LSERK4(plan, 0.1, fused=True)._fused_adjoint
    assert not LSERK4(plan, 0.1, fused=False)._fused_adjoint


def test_maxwell_p4_keeps_the_unfused_reverse_stage():
    """Maxwell P4's fused reverse step measured slower (0.87x): the default
    is the unfused pair, while wave P3 fuses it."""
    assert not LSERK4(_mock_plan("maxwell", 4), 0.1)._fused_adjoint
    assert LSERK4(_mock_plan("wave", 3), 0.1)._fused_adjoint


def test_euler_reverse_stage_is_never_fused():
    assert not LSERK4(_mock_plan("euler", 2), 0.1, fused=True)._fused_adjoint
