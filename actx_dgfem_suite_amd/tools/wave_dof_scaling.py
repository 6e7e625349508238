"""Wave-operator throughput against problem size, mirroring
tools/wave_dof_scaling.py of the reference: for each polynomial degree and
each target size (DOFs per field), build the wave right-hand side with
``get_wave_rhs``, time it with the runner's protocol (up to 5 warm-up calls or
2 s, then blocks of 3 calls until 50 calls or 4 s, the queue finished around
every block) and report GDOF/s = actual DOFs per field x (dim + 1) fields /
time per call.

    python -m actx_dgfem_suite_amd.tools.wave_dof_scaling --dim 3 --degrees 1,2,3,4 \\
        [--ndofs 1e6,5e6,1e7,2e7] [--actx hip:dgfem]

The reference's size lists are per device (Titan V, H200 NVL); the default
list here spans 1M-20M DOFs per field, where one MI355X holds every field many
times over.
"""

from __future__ import annotations

import argparse
import gc
from collections.abc import Sequence
from time import time

import numpy as np

DEFAULT_NDOFS = (1_000_000, 5_000_000, 10_000_000, 20_000_000)


def stringify_dofs_per_s(dofs_per_s: float) -> str:
    return "N/A" if np.isnan(dofs_per_s) else f"{dofs_per_s * 1e-9:.3f}"


def time_rhs(compiled, args, finish) -> float:
    """Seconds per call, with the reference runner's warm-up and timing rounds."""
    i_warmup, t_warmup = 0, 0.0
    while i_warmup < 5 and t_warmup < 2:
        finish()
        t0 = time()
        compiled(*args)
        finish()
        t_warmup += time() - t0
        i_warmup += 1
    i_timing, t_rhs = 0, 0.0
    while i_timing < 50 and t_rhs < 4:
        finish()
        t0 = time()
        for _ in range(3):
            compiled(*args)
        finish()
        t_rhs += time() - t0
        i_timing += 3
    return t_rhs / i_timing


def main(dim: int, degrees: Sequence[int], ndofs_list: Sequence[int],
         actx_name: str = "hip:dgfem") -> np.ndarray:
    from tabulate import tabulate

    from actx_dgfem_suite_amd.equations.wave import get_wave_rhs
    from actx_dgfem_suite_amd.measure import finish_command_queue, instantiate_actx_t
    from actx_dgfem_suite_amd.run.__main__ import _NAME_TO_ACTX_CLASS
    from actx_dgfem_suite_amd.utils import (
        count_nodes_regular_rect_mesh,
        get_ndof_for_regular_rect_mesh,
        get_nel_1d_for_regular_rect_mesh,
    )

    actx_t = _NAME_TO_ACTX_CLASS[actx_name]
    throughput = np.empty([len(degrees), len(ndofs_list)])
    for ideg, degree in enumerate(degrees):
        for ind, ndofs in enumerate(ndofs_list):
            gc.collect()
            actx = instantiate_actx_t(actx_t)
            rhs, args = get_wave_rhs(actx=actx, dim=dim, order=degree, ndofs=int(ndofs))
            compiled = actx.compile(rhs)
            t = time_rhs(compiled, args, lambda: finish_command_queue(actx))
            nel_1d = get_nel_1d_for_regular_rect_mesh(dim, degree, int(ndofs))
            # the reference's count (its 3D P1 figure counts 6 nodes per tet),
            # so the table compares with the reference tool's
            actual = get_ndof_for_regular_rect_mesh(dim, degree, nel_1d)
            nodes = count_nodes_regular_rect_mesh(dim, degree, nel_1d)
            throughput[ideg, ind] = actual * (dim + 1) / t
            print(f"P{degree} target {ndofs / 1e6:g}M: actual {actual / 1e6:.3f}M DOFs/field "
                  f"({nodes / 1e6:.3f}M nodes/field, {nodes * (dim + 1) / t / 1e9:.3f} GDOF/s "
                  f"on nodes), {1e3 * t:.3f} ms per RHS", flush=True)
            del rhs, args, compiled, actx
            gc.collect()
    print(f"GDOFS/s for {dim}D-wave ({actx_name}):")
    table = [[f"P{degree}", *[stringify_dofs_per_s(throughput[i, j])
                              for j in range(len(ndofs_list))]]
             for i, degree in enumerate(degrees)]
    print(tabulate(table, tablefmt="fancy_grid",
                   headers=[""] + [f"{n / 1e6:g}M" for n in ndofs_list]))
    return throughput


def _parse(argv=None):
    p = argparse.ArgumentParser(prog="python -m actx_dgfem_suite_amd.tools.wave_dof_scaling",
                                description="GDOF/s of the wave operator against problem size")
    p.add_argument("--dim", type=int, default=3)
    p.add_argument("--degrees", type=str, default="1,2,3,4")
    p.add_argument("--ndofs", type=str, default=",".join(str(n) for n in DEFAULT_NDOFS),
                   help="comma separated target DOFs per field")
    p.add_argument("--actx", type=str, default="hip:dgfem")
    return p.parse_args(argv)


if __name__ == "__main__":
    a = _parse()
    main(dim=a.dim, degrees=[int(k) for k in a.degrees.split(",")],
         ndofs_list=[int(float(k)) for k in a.ndofs.split(",")], actx_name=a.actx)
