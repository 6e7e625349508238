"""Benchmark runner mirroring run/__main__.py of the reference.

    python -m actx_dgfem_suite_amd.run --equations wave,euler --dims 2,3 \\
        --degrees 1,2,3,4 --actxs hip:dgfem,hip:dgfem_lds

For every (array context, dim, equation, degree) it builds the right-hand side
with ``get_rhs`` at the reference's size tiers, compiles it with the array
context, optionally checks the value, times it with the reference's protocol
(run/__main__.py:101-135: up to 5 warm-up calls or 2 s, then blocks of 3
calls with a queue finish around each block until 50 calls or 4 s) and prints
GFLOP/s tables with the roofline rate (perf_analysis rule, restated in
``actx_dgfem_suite_amd.perf_analysis``).

Value check: the reference compares against its ``NumpyArrayContext`` at
rtol = atol = 1e-5 (run/__main__.py:77-96), i.e. the same operator code
evaluated eagerly with numpy.  Here, on tetrahedra, the check traces the
operator as written against the grudge-shaped layer (``lazy.operators``:
fluxes in Python, grudge's unfused einsums) and evaluates that DAG on the
host with numpy (``lazy.host_eval``), which shares no code with the HIP
kernels' flux functors or operator tiles.  The tolerance is the reference's
(1e-3 for the mixed-precision contexts, whose FP32 state rounds at about
1e-7 relative before differentiation).  The lazy operators are written for
tetrahedra, so on triangles (``--dims 2``) the check compares against the
LDS-staged context, which runs the same 2D kernel: there it only confirms a
repeatable result.  The host evaluation is the check's alone; the timed
callables run on the GPU.
"""

from __future__ import annotations

import argparse
import gc
from collections.abc import Sequence
from time import time

import numpy as np

from actx_dgfem_suite_amd.arraycontext import (
    HIPDGFEMArrayContext,
    HIPDGFEMBF16x3ArrayContext,
    HIPDGFEMFP32ArrayContext,
    HIPDGFEMLDSArrayContext,
)
from actx_dgfem_suite_amd.containers import flatten_fields
from actx_dgfem_suite_amd.lazy import HIPDGFEMLazyArrayContext
from actx_dgfem_suite_amd.lazy.host_eval import evaluate
from actx_dgfem_suite_amd.measure import finish_command_queue, instantiate_actx_t
from actx_dgfem_suite_amd.perf_analysis import get_float64_flops, get_roofline_flop_rate
from actx_dgfem_suite_amd.rhs_builder import get_rhs

# the plug-in table (run/__main__.py:209-216)
_NAME_TO_ACTX_CLASS = {
    "hip:dgfem": HIPDGFEMArrayContext,
    "hip:dgfem_lds": HIPDGFEMLDSArrayContext,
    "hip:dgfem_fp32": HIPDGFEMFP32ArrayContext,
    "hip:dgfem_bf16x3": HIPDGFEMBF16x3ArrayContext,
    # traces the operator into an einsum DAG, normalises and matches it
    "hip:dgfem_lazy": HIPDGFEMLazyArrayContext,
}
_ACTX_CLASS_TO_NAME = {v: k for k, v in _NAME_TO_ACTX_CLASS.items()}


def stringify_flops(flops: float) -> str:
    return "N/A" if np.isnan(flops) else f"{flops * 1e-9:.1f}"


def _host_value(equation: str, dim: int, degree: int) -> list:
    """The traced operator of ``equation`` evaluated on the host with numpy,
    on the state ``get_rhs`` builds for it."""
    lazy = HIPDGFEMLazyArrayContext()
    rhs, args = get_rhs(equation, lazy, dim, degree)
    pos, phs, _, leaves = lazy.trace(rhs, args)
    state = [t.detach().cpu().numpy() for t in flatten_fields(args[pos])]
    return evaluate(leaves, {ph.name: s for ph, s in zip(phs, state)})


def _verify(actx_t, equation: str, dim: int, degree: int, output, actx) -> None:
    got_fields = flatten_fields(actx.to_numpy(output))
    if dim == 3:
        want_fields = _host_value(equation, dim, degree)
    else:
        check_t = HIPDGFEMLDSArrayContext if actx_t is not HIPDGFEMLDSArrayContext \
            else HIPDGFEMArrayContext
        other = instantiate_actx_t(check_t)
        rhs, args = get_rhs(equation, other, dim, degree)
        want_fields = flatten_fields(other.to_numpy(other.compile(rhs)(*args)))
        del other, rhs, args
    assert len(got_fields) == len(want_fields)
    for got, want in zip(got_fields, want_fields):
        np.testing.assert_allclose(np.asarray(got, dtype=np.float64), want,
                                   rtol=actx_t.verify_rtol, atol=actx_t.verify_rtol)
    del want_fields
    gc.collect()


def _get_flop_rate(actx_t, equation: str, dim: int, degree: int, verify: bool) -> float:
    actx = instantiate_actx_t(actx_t)
    rhs, args = get_rhs(equation, actx, dim, degree)
    compiled_rhs = actx.compile(rhs)

    output = compiled_rhs(*args)
    if verify:
        _verify(actx_t, equation, dim, degree, output, actx)
    del output

    i_warmup, t_warmup = 0, 0.0
    while i_warmup < 5 and t_warmup < 2:
        finish_command_queue(actx)
        t_start = time()
        compiled_rhs(*args)
        finish_command_queue(actx)
        t_warmup += time() - t_start
        i_warmup += 1

    i_timing, t_rhs = 0, 0.0
    while i_timing < 50 and t_rhs < 4:
        finish_command_queue(actx)
        t_start = time()
        for _ in range(3):
            compiled_rhs(*args)
        finish_command_queue(actx)
        t_rhs += time() - t_start
        i_timing += 3

    return get_float64_flops(equation, dim, degree) * i_timing / t_rhs


def main(equations: Sequence[str], dims: Sequence[int], degrees: Sequence[int],
         actx_ts: Sequence[type], verify: bool) -> np.ndarray:
    from tabulate import tabulate

    flop_rate = np.empty([len(actx_ts), len(dims), len(equations), len(degrees)])
    roofline = np.empty([len(dims), len(equations), len(degrees)])
    for idim, dim in enumerate(dims):
        for ieq, eq in enumerate(equations):
            for ideg, deg in enumerate(degrees):
                roofline[idim, ieq, ideg] = get_roofline_flop_rate(eq, dim, deg)
    for iactx, actx_t in enumerate(actx_ts):
        for idim, dim in enumerate(dims):
            for ieq, eq in enumerate(equations):
                for ideg, deg in enumerate(degrees):
                    flop_rate[iactx, idim, ieq, ideg] = _get_flop_rate(
                        actx_t, eq, dim, deg, verify=verify)
                    gc.collect()

    for idim, dim in enumerate(dims):
        for ieq, eq in enumerate(equations):
            print(f"GFLOPS/s for {dim}D-{eq}:")
            table = [["", *[_ACTX_CLASS_TO_NAME[a] for a in actx_ts], "Roofline"]]
            for ideg, deg in enumerate(degrees):
                table.append([f"P{deg}",
                              *[stringify_flops(flop_rate[ia, idim, ieq, ideg])
                                for ia in range(len(actx_ts))],
                              stringify_flops(roofline[idim, ieq, ideg])])
            print(tabulate(table, tablefmt="fancy_grid"))
    return flop_rate


def _parse(argv=None):
    parser = argparse.ArgumentParser(prog="python -m actx_dgfem_suite_amd.run",
                                     description="Run DG-FEM benchmarks for array contexts")
    parser.add_argument("--equations", metavar="E", type=str, required=True,
                        help="comma separated equations, e.g. 'wave,euler,maxwell'; prefix "
                             "'tiny_' for 1K DOFs per field, 'large_' for 20M (default 4M)")
    parser.add_argument("--dims", metavar="D", type=str, required=True,
                        help="comma separated topological dimensions (2, 3)")
    parser.add_argument("--degrees", metavar="G", type=str, required=True,
                        help="comma separated polynomial degrees, e.g. 1,2,3")
    parser.add_argument("--actxs", metavar="A", type=str, required=True,
                        help="comma separated array context names, e.g. "
                             f"'{','.join(_NAME_TO_ACTX_CLASS)}'")
    parser.add_argument("--no-verify", action="store_true", default=False,
                        help="skips value correctness checks")
    return parser.parse_args(argv)


if __name__ == "__main__":
    a = _parse()
    main(equations=[k.strip() for k in a.equations.split(",")],
         dims=[int(k.strip()) for k in a.dims.split(",")],
         degrees=[int(k.strip()) for k in a.degrees.split(",")],
         actx_ts=[_NAME_TO_ACTX_CLASS[k.strip()] for k in a.actxs.split(",")],
         verify=not a.no_verify)
