"""CPU tests of the FLOP / footprint model, the peak table and the runner's
command line (no GPU)."""

import itertools

import numpy as np
import pytest

from actx_dgfem_suite_amd.consts import lookup_device
from actx_dgfem_suite_amd.perf_analysis import (
    box_mesh_counts,
    einsum_flops,
    get_float64_flops,
    get_footprint_bytes,
    get_roofline_flop_rate,
    mesh_counts,
)
from actx_dgfem_suite_amd.utils import get_nel_1d_for_regular_rect_mesh
from actx_dgfem_suite_amd.workloads import algorithmic_bytes


def _brute_force_einsum(spec, sizes):
    """Enumerate every pairwise order of a 2- or 3-operand einsum and apply
    the rule literally: per step, count iterations; mult always, add when an
    index disappears."""
    lhs, out = spec.split("->")
    ops = lhs.split(",")
    best = None
    for order in itertools.permutations(range(len(ops))):
        cur, total, rest = ops[order[0]], 0, [ops[k] for k in order[1:]]
        for k, nxt in enumerate(rest):
            later = "".join(rest[k + 1:])
            idx = set(cur) | set(nxt)
            keep = {i for i in idx if i in out or i in later}
            n = int(np.prod([sizes[i] for i in idx]))
            total += n * (2 if keep != idx else 1)
            cur = "".join(sorted(keep))
        best = total if best is None else min(best, total)
    return best


@pytest.mark.parametrize("spec", ["xre,rij,xej->ei", "re,rij,ej->ei", "ifj,fe,fej->ei",
                                  "ij,ej->ei"])
def test_einsum_rule_matches_brute_force(spec):
    sizes = {"x": 3, "r": 3, "e": 17, "i": 20, "j": 20, "f": 4}
    if spec == "ifj,fe,fej->ei":
        sizes["j"] = 10
    assert einsum_flops(spec, sizes) == _brute_force_einsum(spec, sizes)


def test_matvec_einsum_is_2mn():
    # one contraction: a mult and an add per (e, i, j)
    assert einsum_flops("ij,ej->ei", {"i": 7, "j": 5, "e": 3}) == 2 * 7 * 5 * 3


def test_tier_sizes_and_footprint():
    m = mesh_counts("wave", 3, 3)
    assert m.k == 6 * 33 ** 3 and m.np == 20 and m.nfp == 10
    assert m.nint + m.nbnd == m.nall
    assert get_footprint_bytes("wave", 3, 3) == algorithmic_bytes(m.k, 20, 10, 4)
    assert mesh_counts("tiny_wave", 3, 1).k == 6 * 4 ** 3


def test_flops_grow_with_order_and_exceed_the_contractions():
    from actx_dgfem_suite_amd.workloads import algorithmic_flops
    for eq, nf in (("wave", 4), ("euler", 5), ("maxwell", 6)):
        for p in (1, 2, 3, 4):
            m = mesh_counts(eq, 3, p)
            f = get_float64_flops(eq, 3, p)
            assert f > algorithmic_flops(eq, m.k, m.np, m.nfp, nf)
            assert f / m.nvol > get_float64_flops(eq, 3, max(p - 1, 1)) / \
                mesh_counts(eq, 3, max(p - 1, 1)).nvol or p == 1


def test_roofline_rate_uses_the_peak_table():
    f = get_float64_flops("wave", 3, 1)
    b = get_footprint_bytes("wave", 3, 1)
    rate = get_roofline_flop_rate("wave", 3, 1, device_name="AMD Instinct MI355X")
    assert rate == pytest.approx(f / max(f / 78.6e12, b / 8.0e12))
    assert np.isnan(get_roofline_flop_rate("wave", 3, 1, device_name="some other device"))
    assert lookup_device("AMD Instinct MI355X OAM") == "AMD Instinct MI355X"
    assert lookup_device("AMD Radeon Graphics", "gfx950:sramecc+:xnack-") == "AMD Instinct MI355X"
    assert lookup_device("AMD Radeon Graphics", "gfx942") is None


def test_unknown_names_raise():
    with pytest.raises(NotImplementedError):
        get_float64_flops("burgers", 3, 1)
    with pytest.raises(NotImplementedError):
        mesh_counts("wave", 1, 1)
    from actx_dgfem_suite_amd.measure import finish_command_queue, instantiate_actx_t
    with pytest.raises(NotImplementedError):
        instantiate_actx_t(object)
    with pytest.raises(NotImplementedError):
        finish_command_queue(object())


def test_runner_command_line():
    from actx_dgfem_suite_amd.run.__main__ import _NAME_TO_ACTX_CLASS, _parse
    a = _parse(["--equations", "wave,tiny_euler", "--dims", "3", "--degrees", "1,2",
                "--actxs", "hip:dgfem,hip:dgfem_lds", "--no-verify"])
    assert a.no_verify and a.equations == "wave,tiny_euler"
    assert set(_NAME_TO_ACTX_CLASS) == {"hip:dgfem", "hip:dgfem_lds", "hip:dgfem_fp32",
                                         "hip:dgfem_bf16x3", "hip:dgfem_lazy"}


def test_dof_scaling_command_line():
    from actx_dgfem_suite_amd.tools.wave_dof_scaling import DEFAULT_NDOFS, _parse
    a = _parse(["--dim", "3", "--degrees", "2,3", "--ndofs", "1e6,4e6"])
    assert a.dim == 3 and a.degrees == "2,3" and a.ndofs == "1e6,4e6"
    assert _parse([]).ndofs == ",".join(str(n) for n in DEFAULT_NDOFS)


@pytest.mark.parametrize("degree,nel", [(1, 3), (2, 4), (4, 2)])
def test_triangle_mesh_counts_match_the_mesh(degree, nel):
    """2D sizes: 2 n^2 triangles, 4 n boundary edges, Np and Nfp of the
    reference triangle -- read off an actual square mesh."""
    from actx_dgfem_suite_amd.discretization.triangle import make_square_discretization
    d = make_square_discretization(degree, nel)
    m = box_mesh_counts(degree, nel, dim=2)
    assert (m.k, m.np, m.nfp, m.nfaces) == (d.nelements, d.np, d.nfp, 3)
    assert m.nbdry == int((d.ext_map[:, ::d.nfp] < 0).sum())
    assert m.nall == d.ext_map.size
    assert m.nbnd == int((d.ext_map < 0).sum())


def test_triangle_footprint_and_flops():
    """2D footprint: 3 or 4 fields in and out, 13 geometry doubles and 3 Nfp
    map entries per triangle; the counting rule's contractions in 2D are the
    einsums with x, r of size 2 and 3 faces."""
    m = box_mesh_counts(3, 5, dim=2)
    assert get_footprint_bytes("tiny_wave", 2, 3) == \
        box_mesh_counts(3, get_nel_1d_for_regular_rect_mesh(2, 3, 1000), dim=2).k * \
        (2 * 3 * 10 * 8 + 13 * 8 + 4 * 3 * 4)
    lift = einsum_flops("ifj,fe,fej->ei", {"i": 10, "f": 3, "j": 4, "e": m.k})
    assert lift == m.k * 3 * 4 * 10 * 2 + m.k * 3 * 4       # scale, then matvec
    for eq in ("tiny_wave", "tiny_euler", "tiny_maxwell"):
        f = [get_float64_flops(eq, 2, deg) for deg in (1, 2, 3, 4)]
        assert all(x > 0 for x in f)
