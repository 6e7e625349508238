"""``get_rhs`` dispatch mirroring rhs_builder.py:9-45 of the reference:
``tiny_`` (1e3 DOFs per field), default (4e6) and ``large_`` (2e7) tiers."""

from __future__ import annotations

from typing import Any

NDOFS_FULL = 4_000_000
NDOFS_TINY = 1_000
NDOFS_LARGE = 20_000_000


def parse_equation(equation: str) -> tuple[int, str]:
    """``(ndofs per field, base equation)`` of a tiered name such as
    ``tiny_wave`` (rhs_builder.py:14-29)."""
    if equation.startswith("tiny_"):
        ndofs, base = NDOFS_TINY, equation.removeprefix("tiny_")
    elif equation.startswith("large_"):
        ndofs, base = NDOFS_LARGE, equation.removeprefix("large_")
    else:
        ndofs, base = NDOFS_FULL, equation
    if base not in ("wave", "euler", "maxwell"):
        raise NotImplementedError(f"equation '{equation}' not supported")
    return ndofs, base


def get_rhs(equation: str, actx, dim: int, degree: int) -> tuple[Any, tuple[Any, ...]]:
    """Return ``(rhs_callable, call_args)`` for ``equation``.  On the tracing
    context the callable is the operator written against the grudge-shaped
    layer (``lazy.operators``), which the context traces and matches."""
    ndofs, base = parse_equation(equation)
    from actx_dgfem_suite_amd.lazy import HIPDGFEMLazyArrayContext
    if isinstance(actx, HIPDGFEMLazyArrayContext):
        from actx_dgfem_suite_amd.lazy import operators as lazy_ops
        builder = {"wave": lazy_ops.get_wave_rhs, "euler": lazy_ops.get_euler_rhs,
                   "maxwell": lazy_ops.get_maxwell_rhs}[base]
        return builder(actx=actx, dim=dim, order=degree, ndofs=ndofs)
    if base == "wave":
        from actx_dgfem_suite_amd.equations.wave import get_wave_rhs
        return get_wave_rhs(actx=actx, dim=dim, order=degree, ndofs=ndofs)
    if base == "euler":
        from actx_dgfem_suite_amd.equations.euler import get_euler_rhs
        return get_euler_rhs(actx=actx, dim=dim, order=degree, ndofs=ndofs)
    if base == "maxwell":
        from actx_dgfem_suite_amd.equations.maxwell import get_maxwell_rhs
        return get_maxwell_rhs(actx=actx, dim=dim, order=degree, ndofs=ndofs)
    raise NotImplementedError(f"equation '{equation}' not supported")
