"""FLOP and footprint model of the DG right-hand sides, and the roofline rate.

``get_float64_flops`` restates the reference's counting rule
(perf_analysis.py:249-417, 533-576):

* every elementwise binary operation (add, sub, mult, div), ``where`` and the
  C99 calls ``sqrt``/``max``/``min`` cost 1 flop per output element;
* an einsum costs, along its optimal pairwise contraction path, one mult per
  iteration of each pairwise step, plus one add per iteration when that step
  contracts an index (perf_analysis.py:366-399; the optimal path over <= 3
  operands is found here by brute force, which is what ``opt_einsum``'s
  "optimal" does for so few operands);
* folded constants (operators, geometry) cost nothing.

The reference applies the rule to the pytato DAG of the optimised operator,
which this build cannot construct (pytato is not installed).  The DAG is
therefore restated by hand from the operator code: the wave operator line by
line from equations/wave.py:80-200, Euler and Maxwell from the textbook
operators that upstream grudge implements.  The counts are parity-unpinned
against the reference's integers.

``get_footprint_bytes`` is the fused kernel's algorithmic footprint (state in
and out once, per-element geometry, the exterior-node map), which replaces
the reference's sum over the loopy sub-kernels' arguments
(perf_analysis.py:580-620): one fused kernel has one argument set.
"""

from __future__ import annotations

import itertools
from dataclasses import dataclass
from functools import cache

from actx_dgfem_suite_amd.discretization.refelem import nfp_for_order, np_for_order
from actx_dgfem_suite_amd.rhs_builder import parse_equation
from actx_dgfem_suite_amd.utils import get_nel_1d_for_regular_rect_mesh
from actx_dgfem_suite_amd.workloads import algorithmic_bytes


def einsum_flops(spec: str, sizes: dict[str, int]) -> int:
    """Flops of ``spec`` (e.g. ``"xre,rij,xej->ei"``) under the reference
    rule, along the cheapest pairwise contraction order."""
    lhs, out = spec.split("->")
    operands = lhs.split(",")

    def pair_cost(a: str, b: str, keep: str) -> tuple[int, str]:
        idx = set(a) | set(b)
        result = "".join(sorted(i for i in idx if i in keep))
        iters = 1
        for i in idx:
            iters *= sizes[i]
        contracted = any(i not in keep for i in idx)
        return iters * (2 if contracted else 1), result

    def best(ops: list[str]) -> int:
        if len(ops) == 1:
            return 0
        costs = []
        for i, j in itertools.combinations(range(len(ops)), 2):
            rest = [o for k, o in enumerate(ops) if k not in (i, j)]
            keep = set(out) | set("".join(rest))
            c, res = pair_cost(ops[i], ops[j], "".join(keep))
            costs.append(c + best(rest + [res]))
        return min(costs)

    return best(operands)


@dataclass(frozen=True)
class MeshCounts:
    k: int          # elements
    np: int         # nodes per element
    nfp: int        # nodes per face
    nbdry: int      # boundary faces

    @property
    def nvol(self) -> int:
        return self.k * self.np

    @property
    def nint(self) -> int:
        """Interior trace nodes (both orientations of every interior face)."""
        return (4 * self.k - self.nbdry) * self.nfp

    @property
    def nbnd(self) -> int:
        return self.nbdry * self.nfp

    @property
    def nall(self) -> int:
        return 4 * self.k * self.nfp


def box_mesh_counts(degree: int, nel_1d: int) -> MeshCounts:
    """Sizes of the box mesh of ``nel_1d``^3 hexahedra x 6 tetrahedra."""
    n = int(nel_1d)
    return MeshCounts(k=6 * n ** 3, np=np_for_order(degree), nfp=nfp_for_order(degree),
                      nbdry=12 * n * n)


def mesh_counts(equation: str, dim: int, degree: int) -> MeshCounts:
    """Sizes of the reference tier's box mesh (rhs_builder.py:9-45)."""
    if dim != 3:
        raise NotImplementedError("only 3D operators are built")
    ndofs, _ = parse_equation(equation)
    return box_mesh_counts(degree, get_nel_1d_for_regular_rect_mesh(dim, degree, ndofs))


def float64_flops_on_mesh(equation: str, m: MeshCounts) -> int:
    """The counting rule applied to ``equation`` on the mesh sizes ``m``."""
    _, base = parse_equation(equation)
    return {"wave": _wave_flops, "euler": _euler_flops, "maxwell": _maxwell_flops}[base](m)


def _contractions(m: MeshCounts, n_div: int, n_grad: int, n_lift: int) -> int:
    sz = {"x": 3, "r": 3, "e": m.k, "i": m.np, "j": m.np}
    div = einsum_flops("xre,rij,xej->ei", sz)
    grad = einsum_flops("re,rij,ej->ei", sz)
    lift = einsum_flops("ifj,fe,fej->ei", {"i": m.np, "f": 4, "j": m.nfp, "e": m.k})
    return n_div * div + n_grad * grad + n_lift * lift


def _wave_flops(m: MeshCounts) -> int:
    # equations/wave.py:80-114, per trace node of a trace pair
    per_trace = (2 + 6          # u.avg, v.avg: (int + ext) * 0.5
                 + 1 + 3        # u.diff, v.diff
                 + 5 + 3        # v.avg @ n, u.avg * n
                 + 5            # v.diff @ n
                 + 1 + 1 + 3    # 0.5 * u.diff, 0.5 * v_jump, * n
                 + 4            # flux_weak += ...
                 + 4)           # c * flux_weak
    flops = per_trace * (m.nint + m.nbnd)
    flops += m.nbnd              # dir_bc u = -dir_u (:155)
    flops += 4 * m.nall          # boundary + interior fluxes summed on all_faces
    # -c * weak_div(v), -c * weak_grad(u) (:187-190); face_mass; the sum
    flops += _contractions(m, n_div=1, n_grad=3, n_lift=4)
    flops += 4 * m.nvol          # -c * (...)
    flops += 4 * m.nvol          # volume + surface
    return flops


def _euler_side(with_speed: bool) -> int:
    # u = m / rho (3); |u|^2 (5); p = (g-1)(E - 0.5 rho |u|^2) (4)
    f = 3 + 5 + 4
    # F.n: m.n (5); (E + p) u.n (5 + 2); m (u.n) + p n (9)
    f += 5 + 7 + 9
    if with_speed:
        f += 1 + 3 + 1           # |u| = sqrt, c = sqrt(g p / rho), |u| + c
    return f


def _euler_flops(m: MeshCounts) -> int:
    # volume: u (3), p (9 with |u|^2), flux tensor (E + p) u (4), m u^T + p I (12)
    per_vol = 3 + 9 + 4 + 12
    flops = per_vol * m.nvol
    # LF flux per trace node: both sides, max, averages, jump, dissipation
    per_trace = 2 * _euler_side(True) + 1 + 10 + 5 + (1 + 5) + 5
    flops += per_trace * (m.nint + m.nbnd)
    flops += 12 * m.nbnd         # wall state m - 2 (m.n) n
    flops += 5 * m.nall          # boundary + interior fluxes summed
    flops += _contractions(m, n_div=5, n_grad=0, n_lift=5)
    flops += 5 * m.nvol          # volume - surface
    return flops


def _maxwell_flops(m: MeshCounts) -> int:
    # local curl: 18 reference-to-physical derivatives (strong D), 6 differences,
    # 3 negations for -curl E
    flops = _contractions(m, n_div=0, n_grad=18, n_lift=6) + 9 * m.nvol
    # flux per trace node: jumps (6), n x [H], n x [E] (18), n x (n x .) (18),
    # scalings (12), sums (6)
    per_trace = 6 + 18 + 18 + 12 + 6
    flops += per_trace * (m.nint + m.nbnd)
    flops += 3 * m.nbnd          # PEC: E+ = -E-
    flops += 6 * m.nall          # boundary + interior fluxes summed
    flops += 6 * m.nvol          # volume - surface
    return flops


@cache
def get_float64_flops(equation: str, dim: int, degree: int) -> int:
    return float64_flops_on_mesh(equation, mesh_counts(equation, dim, degree))


@cache
def get_footprint_bytes(equation: str, dim: int, degree: int) -> int:
    m = mesh_counts(equation, dim, degree)
    _, base = parse_equation(equation)
    nf = {"wave": 4, "euler": 5, "maxwell": 6}[base]
    return algorithmic_bytes(m.k, m.np, m.nfp, nf)


def get_roofline_flop_rate(equation: str, dim: int, degree: int,
                           roofline_model: str = "libparanumal:global_ai",
                           device_name: str | None = None) -> float:
    """FLOP rate at the roofline ``t = max(F / F_peak, B / BW_peak)``
    (perf_analysis.py:623-653); NaN for a device not in the peak table."""
    from actx_dgfem_suite_amd.consts import (
        DEV_TO_PEAK_BW,
        DEV_TO_PEAK_F64_GFLOPS,
        lookup_device,
    )
    if roofline_model != "libparanumal:global_ai":
        raise NotImplementedError(f"unknown roofline model: {roofline_model}")
    arch = ""
    if device_name is None:
        import torch
        device_name = ""
        if torch.cuda.is_available():
            device_name = torch.cuda.get_device_name(0)
            arch = getattr(torch.cuda.get_device_properties(0), "gcnArchName", "")
    key = lookup_device(device_name, arch)
    if key is None:
        return float("nan")
    nflops = get_float64_flops(equation, dim, degree)
    nbytes = get_footprint_bytes(equation, dim, degree)
    t = max(nflops * 1e-9 / DEV_TO_PEAK_F64_GFLOPS[key], nbytes * 1e-9 / DEV_TO_PEAK_BW[key])
    return nflops / t
