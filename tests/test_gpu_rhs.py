"""GPU parity: the HIP right-hand sides against the oracle (small meshes),
against the golden fixtures, and through size-independent properties at the
full benchmark sizes.  All calls go through the C ABI (``DGPlan``)."""

import glob
import os

import numpy as np
import pytest
import torch

from actx_dgfem_suite_amd.discretization import make_box_discretization
from actx_dgfem_suite_amd.plan import DGPlan
from actx_dgfem_suite_amd.workloads import WORKLOADS, initial_state
from oracle import dg_oracle
from tests._helpers import (
    discr_pair,
    discr_pair_2d,
    field_errors,
    oracle_rhs,
    random_state,
    rel_l2,
)

pytestmark = pytest.mark.gpu

PARAM = {"wave": 1.0, "euler": 1.4, "maxwell": 0.5}
# FP64 drop-in form: relative L2 per field against the FP64 oracle
TOL_FP64 = 1e-12


def _dev32(a, device):
    return [torch.from_numpy(np.ascontiguousarray(f)).to(device).float() for f in a]


def _run(plan, q_np, device):
    q = [torch.from_numpy(np.ascontiguousarray(f)).to(device) for f in q_np]
    out = plan.rhs(q)
    torch.cuda.synchronize()
    return np.stack([t.cpu().numpy() for t in out])


@pytest.mark.parametrize("variant", ["mfma", "lds"])
@pytest.mark.parametrize("equation", ["wave", "euler", "maxwell"])
@pytest.mark.parametrize("order", [1, 2, 3, 4])
def test_rhs_matches_oracle(equation, order, variant, device):
    print(f"[case] rhs {variant} {equation} P{order} n=3", flush=True)
    d, od = discr_pair(order, 3)
    q = random_state(equation, od.x, seed=order)
    plan = DGPlan(d, equation, PARAM[equation], variant=variant)
    got = _run(plan, q, device)
    ref = oracle_rhs(equation, od, q, PARAM[equation])
    errs = field_errors(got, ref)
    print(f"  rel L2 per field: {['%.2e' % e for e in errs]}", flush=True)
    assert max(errs) < TOL_FP64, (equation, order, errs)


@pytest.mark.parametrize("variant", ["mfma", "lds"])
@pytest.mark.parametrize("equation,order", [("wave", 3), ("euler", 2), ("maxwell", 4)])
def test_rhs_matches_oracle_many_batches(equation, order, variant, device):
    """A mesh large enough that every persistent block runs several batches."""
    print(f"[case] rhs {variant} {equation} P{order} n=12", flush=True)
    d, od = discr_pair(order, 12)
    q = random_state(equation, od.x, seed=7)
    plan = DGPlan(d, equation, PARAM[equation], variant=variant)
    got = _run(plan, q, device)
    ref = oracle_rhs(equation, od, q, PARAM[equation])
    assert max(field_errors(got, ref)) < TOL_FP64


GOLDEN = sorted(glob.glob(os.path.join(os.path.dirname(__file__), "golden", "*.npz")))


@pytest.mark.parametrize("path", GOLDEN, ids=[os.path.basename(p) for p in GOLDEN])
def test_rhs_matches_golden(path, device):
    print(f"[case] golden {os.path.basename(path)}", flush=True)
    g = np.load(path, allow_pickle=False)
    tri = int(g["dim"]) == 2 if "dim" in g.files else False
    if tri:
        from actx_dgfem_suite_amd.discretization.triangle import make_square_discretization
        d = make_square_discretization(int(g["order"]), int(g["nel_1d"]))
    else:
        d = make_box_discretization(int(g["order"]), int(g["nel_1d"]))
    plan = DGPlan(d, str(g["equation"]), float(g["param"]))
    got = _run(plan, g["q"], device)
    assert max(field_errors(got, g["rhs"])) < TOL_FP64


# {{{ full-size properties

def _poly_fields(x, order, nf, seed):
    rng = np.random.default_rng(seed)
    exps = [(a, b, c) for a in range(order + 1) for b in range(order + 1 - a)
            for c in range(order + 1 - a - b)]
    vals, grads = [], []
    for _ in range(nf):
        coef = rng.standard_normal(len(exps))
        v = np.zeros(x.shape[1:])
        g = [np.zeros(x.shape[1:]) for _ in range(3)]
        for cf, e in zip(coef, exps):
            v += cf * x[0] ** e[0] * x[1] ** e[1] * x[2] ** e[2]
            for dd in range(3):
                if e[dd]:
                    ee = list(e)
                    ee[dd] -= 1
                    g[dd] += cf * e[dd] * x[0] ** ee[0] * x[1] ** ee[1] * x[2] ** ee[2]
        vals.append(v)
        grads.append(g)
    return np.stack(vals), grads


def _interior(d):
    return np.nonzero(np.all(d.connectivity.nbr_elem >= 0, axis=1))[0]


@pytest.mark.parametrize("name", ["wave_p3", "maxwell_p4"])
def test_full_size_polynomial_exactness(name, device):
    """At the benchmark sizes: continuous polynomial states of degree N give
    the exact derivatives on every interior element (no oracle needed)."""
    w = WORKLOADS[name]
    print(f"[case] full-size polynomial {name}", flush=True)
    d = w.discretization()
    nf = 4 if w.equation == "wave" else 6
    q, g = _poly_fields(d.nodes, w.order, nf, seed=5)
    plan = DGPlan(d, w.equation, w.param)
    got = _run(plan, q, device)
    inner = _interior(d)
    if w.equation == "wave":
        exact = [sum(g[1 + k][k] for k in range(3)), *(g[0][k] for k in range(3))]
    else:
        def curl(gg):
            return [gg[2][1] - gg[1][2], gg[0][2] - gg[2][0], gg[1][0] - gg[0][1]]
        ce, ch = curl(g[0:3]), curl(g[3:6])
        exact = [*ch, *(-c for c in ce)]
    for c in range(nf):
        scale = np.abs(exact[c][inner]).max()
        err = np.abs(got[c][inner] - exact[c][inner]).max()
        assert err < 1e-10 * max(scale, 1.0), (c, err, scale)


@pytest.mark.parametrize("name", ["wave_p3", "euler_p2", "maxwell_p4"])
def test_full_size_linearity_or_consistency(name, device):
    """Wave/Maxwell: L(a x + b y) = a L(x) + b L(y).  Euler: the state at rest
    is steady and mass and energy are conserved with walls."""
    w = WORKLOADS[name]
    print(f"[case] full-size invariants {name}", flush=True)
    d = w.discretization()
    plan = DGPlan(d, w.equation, w.param)
    q0 = initial_state(w, d)
    if w.equation == "euler":
        shape = q0.shape[1:]
        rest = np.stack([np.full(shape, 1.2), np.full(shape, 2.5), *(np.zeros(shape),) * 3])
        r = _run(plan, rest, device)
        # volume and face terms are each ~ p * Fscale * |LIFT| and must cancel
        term = 1.0 * d.geometry.fscale.max() * np.abs(d.ref.lift).max()
        assert np.abs(r).max() < 1e-13 * term
        rhs = _run(plan, q0, device)
        wts = d.geometry.jac[:, None] * (np.ones(d.np) @ d.ref.mass)[None, :]
        for c in (0, 1):
            assert abs(np.sum(wts * rhs[c])) < 1e-12 * np.sum(wts * np.abs(rhs[c]))
        return
    rng = np.random.default_rng(11)
    y = rng.standard_normal(q0.shape)
    a, b = 0.75, -1.25
    lx = _run(plan, q0, device)
    ly = _run(plan, y, device)
    lxy = _run(plan, a * q0 + b * y, device)
    assert rel_l2(lxy, a * lx + b * ly) < 1e-13


def test_tiny_wave_workload_against_oracle(device):
    """The CPU-runnable workload: 4^3 x 6 tets, P1, one RHS."""
    w = WORKLOADS["tiny_wave_p1"]
    d, od = discr_pair(1, 4)
    q = initial_state(w, d)
    plan = DGPlan(d, "wave", 1.0)
    got = _run(plan, q, device)
    ref = oracle_rhs("wave", od, q, 1.0)
    assert got.shape == ref.shape == (4, 384, 4)
    assert max(field_errors(got, ref)) < TOL_FP64

# }}}


# {{{ LSERK4 and error paths

def test_lserk_stage_and_step_match_oracle(device):
    print("[case] lserk", flush=True)
    d, od = discr_pair(2, 3)
    q = random_state("wave", od.x, seed=3)
    plan = DGPlan(d, "wave", 1.0)
    dt = 1e-3
    qt = [torch.from_numpy(f.copy()).to(device) for f in q]
    res = [torch.zeros_like(t) for t in qt]
    rhs = plan.empty_fields()
    for a, b in zip(dg_oracle.LSRK4_A, dg_oracle.LSRK4_B):
        plan.rhs(qt, rhs)
        plan.lserk_stage(a, b, dt, qt, res, rhs)
    torch.cuda.synchronize()
    got = np.stack([t.cpu().numpy() for t in qt])
    ref = dg_oracle.lsrk4_step(lambda t, s: oracle_rhs("wave", od, s, 1.0), q, dt)
    assert rel_l2(got, ref) < 1e-13


@pytest.mark.parametrize("equation", ["wave", "euler", "maxwell"])
@pytest.mark.parametrize("order", [1, 2, 3, 4])
def test_fused_lserk_stage_matches_rhs_then_stage(equation, order, device):
    """dg_rhs_lserk_stage (the stage applied in the MFMA kernel's output)
    against dg_rhs followed by dg_lserk_stage, over the five stages of a step
    on a mesh with a partial last tile; and the fused LSERK4 stepper against
    the oracle's steps: one ``step`` (copy back after five stages) and a
    two-step ``run`` (ten stages, no copy)."""
    print(f"[case] fused lserk {equation} P{order} n=3", flush=True)
    d, od = discr_pair(order, 3)
    param = {"wave": 1.0, "euler": 1.4, "maxwell": 0.5}[equation]
    plan = DGPlan(d, equation, param, variant="mfma")
    assert plan.fused_stage
    q0 = random_state(equation, od.x, seed=order)
    dt = 1e-3
    qa = [torch.from_numpy(f.copy()).to(device) for f in q0]
    qb = [t.clone() for t in qa]
    qc = [torch.empty_like(t) for t in qa]
    ra = [torch.full_like(t, float("nan")) for t in qa]   # a == 0 must not read res
    rb = [t.clone() for t in ra]
    rhs = plan.empty_fields()
    for a, b in zip(dg_oracle.LSRK4_A, dg_oracle.LSRK4_B):
        plan.rhs(qa, rhs)
        plan.lserk_stage(a, b, dt, qa, ra, rhs)
        plan.rhs_lserk_stage(a, b, dt, qb, qc, rb)
        qb, qc = qc, qb
    torch.cuda.synchronize()
    got, ref = np.stack([t.cpu().numpy() for t in qb]), np.stack([t.cpu().numpy() for t in qa])
    assert rel_l2(got, ref) < 1e-14
    from actx_dgfem_suite_amd.timestepping import LSERK4
    st = LSERK4(plan, dt, fused=True)
    assert st.fused
    qs = [torch.from_numpy(f.copy()).to(device) for f in q0]
    st.step(qs)
    torch.cuda.synchronize()
    step = np.stack([t.cpu().numpy() for t in qs])
    o1 = dg_oracle.lsrk4_step(lambda t, s: oracle_rhs(equation, od, s, param), q0, dt)
    assert rel_l2(step, o1) < 1e-13
    qs = [torch.from_numpy(f.copy()).to(device) for f in q0]
    st.run(qs, 2)
    torch.cuda.synchronize()
    two = np.stack([t.cpu().numpy() for t in qs])
    o2 = dg_oracle.lsrk4_step(lambda t, s: oracle_rhs(equation, od, s, param), o1, dt)
    assert rel_l2(two, o2) < 1e-13


def test_fused_lserk_stage_rejects_misuse(device):
    d, _ = discr_pair(2, 2)
    plan = DGPlan(d, "wave", 1.0, variant="mfma")
    q = plan.empty_fields()
    with pytest.raises(ValueError, match="q_out must not overlap q_in"):
        plan.rhs_lserk_stage(0.0, 0.5, 1e-3, q, q, plan.empty_fields())
    q2 = plan.empty_fields()
    with pytest.raises(ValueError, match="res must not overlap q_in"):
        plan.rhs_lserk_stage(0.0, 0.5, 1e-3, q, q2, q)
    with pytest.raises(ValueError, match="res must not overlap q_out"):
        plan.rhs_lserk_stage(0.0, 0.5, 1e-3, q, q2, q2)
    # a partial overlap: q_out[0] starts half-way into q_in[1]
    n = q[0].numel()
    big = torch.empty(2 * n, dtype=torch.float64, device=device)
    qa = [q[0], big[:n].view(plan.shape), q[2], q[3]]
    qb = [big[n // 2:n // 2 + n].view(plan.shape), *plan.empty_fields()[1:]]
    with pytest.raises(ValueError, match=r"q_out must not overlap q_in \(q_out\[0\], q_in\[1\]\)"):
        plan.rhs_lserk_stage(0.0, 0.5, 1e-3, qa, qb, plan.empty_fields())
    lds = DGPlan(d, "wave", 1.0, variant="lds")
    assert not lds.fused_stage
    with pytest.raises(NotImplementedError, match="MFMA"):
        lds.rhs_lserk_stage(0.0, 0.5, 1e-3, q, plan.empty_fields(), plan.empty_fields())
    from actx_dgfem_suite_amd.timestepping import LSERK4
    with pytest.raises(ValueError, match="fused"):
        LSERK4(lds, 1e-3, fused=True)
    d2, _ = discr_pair_2d(2, 2)
    tri = DGPlan(d2, "wave", 1.0)
    with pytest.raises(ValueError, match="fused"):
        LSERK4(tri, 1e-3, precision="fp32", fused=True)
    # the default follows the measured table: wave P1 fused, wave P2 not
    d1, _ = discr_pair(1, 2)
    p1 = DGPlan(d1, "wave", 1.0, variant="mfma")
    assert LSERK4(p1, 1e-3).fused and not LSERK4(lds, 1e-3).fused
    assert plan.fused_stage and not LSERK4(plan, 1e-3).fused
    assert not LSERK4(p1, 1e-3, fused=False).fused
    # Maxwell P4 defaults to unfused stages (measured faster there)
    d4, _ = discr_pair(4, 2)
    mx4 = DGPlan(d4, "maxwell", 0.5, variant="mfma")
    assert mx4.fused_stage and not LSERK4(mx4, 1e-3).fused


@pytest.mark.parametrize("precision", ["fp32", "bf16x3"])
@pytest.mark.parametrize("equation", ["wave", "euler", "maxwell"])
@pytest.mark.parametrize("order", [1, 2, 3, 4])
def test_fused_mixed_stage_matches_unfused(precision, equation, order, device):
    """dg_rhs_lserk_stage_mixed against dg_rhs_mixed followed by
    dg_lserk_stage_mixed over the five stages of a step, on a mesh with a
    partial last tile: both apply the stage to the same FP32 RHS, so they
    differ by the rounding of the FP64 update only; the new FP32 shadow is
    the FP64 state rounded.  And the fused mixed LSERK4 stepper against the
    unfused one."""
    print(f"[case] fused mixed lserk {precision} {equation} P{order} n=3", flush=True)
    d, od = discr_pair(order, 3)
    param = {"wave": 1.0, "euler": 1.4, "maxwell": 0.5}[equation]
    plan = DGPlan(d, equation, param)
    assert plan.fused_stage_mixed
    q0 = random_state(equation, od.x, seed=order)
    dt = 1e-3
    qa = [torch.from_numpy(f.copy()).to(device) for f in q0]
    qb = [t.clone() for t in qa]
    ra = [torch.full_like(t, float("nan")) for t in qa]   # a == 0 must not read res
    rb = [t.clone() for t in ra]
    a32 = [t.float() for t in qa]
    s32, d32 = [t.float() for t in qb], plan.empty_fields(torch.float32)
    rhs32 = plan.empty_fields(torch.float32)
    for a, b in zip(dg_oracle.LSRK4_A, dg_oracle.LSRK4_B):
        plan.rhs_mixed(a32, rhs32, precision)
        plan.lserk_stage_mixed(a, b, dt, qa, ra, rhs32, a32)
        plan.rhs_lserk_stage_mixed(precision, a, b, dt, s32, d32, qb, rb)
        s32, d32 = d32, s32
    torch.cuda.synchronize()
    st = lambda f: np.stack([t.cpu().numpy() for t in f])  # noqa: E731
    assert rel_l2(st(qb), st(qa)) < 1e-14
    assert rel_l2(st(rb), st(ra)) < 1e-12
    assert np.array_equal(st(s32), st(qb).astype(np.float32))
    from actx_dgfem_suite_amd.timestepping import LSERK4
    fused, plain = LSERK4(plan, dt, precision, fused=True), LSERK4(plan, dt, precision, fused=False)
    assert fused.fused and not plain.fused
    q1 = [torch.from_numpy(f.copy()).to(device) for f in q0]
    q2 = [t.clone() for t in q1]
    fused.run(q1, 2)
    plain.run(q2, 2)
    torch.cuda.synchronize()
    assert rel_l2(st(q1), st(q2)) < 1e-14


def test_fused_mixed_stage_rejects_misuse(device):
    d, _ = discr_pair(2, 2)
    plan = DGPlan(d, "wave", 1.0)
    from actx_dgfem_suite_amd.timestepping import LSERK4
    # defaults follow the measured cases: wave P2 fused in BF16x3, not FP32
    assert LSERK4(plan, 1e-3, "bf16x3").fused and not LSERK4(plan, 1e-3, "fp32").fused
    q, res = plan.empty_fields(), plan.empty_fields()
    s32, d32 = plan.empty_fields(torch.float32), plan.empty_fields(torch.float32)
    with pytest.raises(ValueError, match="q32_out must not overlap q32_in"):
        plan.rhs_lserk_stage_mixed("fp32", 0.0, 0.5, 1e-3, s32, s32, q, res)
    with pytest.raises(ValueError, match="res must not overlap q "):
        plan.rhs_lserk_stage_mixed("fp32", 0.0, 0.5, 1e-3, s32, d32, q, q)
    # an FP32 shadow placed inside the FP64 state's memory
    inside = [q[0].view(torch.float32).view(-1)[:q[0].numel()].view(plan.shape), *d32[1:]]
    with pytest.raises(ValueError, match="q32_out must not overlap q"):
        plan.rhs_lserk_stage_mixed("fp32", 0.0, 0.5, 1e-3, s32, inside, q, res)
    with pytest.raises(NotImplementedError):
        plan.rhs_lserk_stage_mixed("fp16", 0.0, 0.5, 1e-3, s32, d32, q, res)
    d2, _ = discr_pair_2d(2, 2)
    tri = DGPlan(d2, "wave", 1.0)
    assert not tri.fused_stage_mixed
    with pytest.raises(NotImplementedError, match="tetrahedral"):
        tri.rhs_lserk_stage_mixed("fp32", 0.0, 0.5, 1e-3, tri.empty_fields(torch.float32),
                                  tri.empty_fields(torch.float32), tri.empty_fields(),
                                  tri.empty_fields())


def test_plan_rejects_bad_fields(device):
    d, _ = discr_pair(1, 2)
    plan = DGPlan(d, "wave", 1.0)
    good = [torch.zeros(plan.shape, dtype=torch.float64, device=device) for _ in range(4)]
    with pytest.raises(TypeError):
        plan.rhs([t.float() for t in good])
    with pytest.raises(ValueError):
        plan.rhs(good[:3])
    with pytest.raises(ValueError):
        plan.rhs([t[:, :2] for t in good])
    with pytest.raises(NotImplementedError):
        plan.rhs(good, precision="fp16")
    # the kernels gather neighbour traces of the input while they write
    with pytest.raises(ValueError, match="fields_out must not overlap fields_in"):
        plan.rhs(good, good)
    with pytest.raises(ValueError, match="overlap"):
        plan.rhs(good, [good[0].clone()] * 4)
    g32 = [t.float() for t in good]
    with pytest.raises(ValueError, match="fields_out must not overlap fields_in"):
        plan.rhs(g32, g32, precision="fp32")

# }}}


@pytest.mark.parametrize("variant", ["mfma", "lds"])
@pytest.mark.parametrize("equation", ["wave", "euler", "maxwell"])
@pytest.mark.parametrize("order", [1, 2, 3, 4])
def test_rhs_single_partial_tile(equation, order, variant, device):
    """A mesh smaller than one 16-element MFMA tile (one hexahedron, 6 tets):
    the padded tile, the tile-ordered tables and the staged state all see
    fewer elements than their slots."""
    print(f"[case] rhs {variant} {equation} P{order} n=1", flush=True)
    d, od = discr_pair(order, 1)
    q = random_state(equation, od.x, seed=13)
    plan = DGPlan(d, equation, PARAM[equation], variant=variant)
    got = _run(plan, q, device)
    ref = oracle_rhs(equation, od, q, PARAM[equation])
    assert max(field_errors(got, ref)) < TOL_FP64
    if variant == "mfma":
        # the mixed form on the same partial tile
        qt = [torch.from_numpy(np.ascontiguousarray(f)).to(device).float() for f in q]
        for prec in ("fp32", "bf16x3"):
            mixed = plan.rhs(qt, precision=prec)
            torch.cuda.synchronize()
            m = np.stack([t.cpu().numpy().astype(np.float64) for t in mixed])
            assert max(field_errors(m, ref)) < 1e-3, prec


@pytest.mark.parametrize("equation,order", [("euler", 1), ("euler", 2), ("euler", 3),
                                            ("maxwell", 2), ("maxwell", 3), ("maxwell", 4),
                                            ("wave", 3)])
def test_mfma_face_geometry_from_rx_or_table(equation, order, device):
    """The MFMA kernel forms its tiles' face geometry from rx where every
    face satisfies Fscale n_f = rx^T c_f, and copies the plan's table where
    one does not (Euler and Maxwell, which take any geometry on this kernel).
    Both match a host evaluation of the same fused forms on their geometry,
    and so do the mixed forms (their kernels form or copy it the same way,
    FP32 Euler and Maxwell on the direct flux); the wave plan on the altered
    geometry goes to the LDS-staged kernel."""
    print(f"[case] face geometry from rx or table {equation} P{order}", flush=True)
    from oracle import dg_cpu
    d = make_box_discretization(order, 3)
    geo = d.packed_geometry().copy()
    geo[7, 9 + 4 * 2 + 3] *= 1.5          # one face's Fscale off the identity

    class Bent:
        def __getattr__(self, name):
            return getattr(d, name)

        def packed_geometry(self):
            return geo

    q = random_state(equation, d.nodes, seed=order)
    for discr in (d, Bent()):
        plan = DGPlan(discr, equation, PARAM[equation], variant="auto")
        expect = "lds" if (equation == "wave" and discr is not d) else "mfma"
        assert plan.kernel_variant == expect
        got = _run(plan, q, device)
        ref = dg_cpu.CpuOperator(discr, equation, PARAM[equation])(q)
        err = max(field_errors(got, ref))
        print(f"  {'bent' if discr is not d else 'affine'}: {plan.kernel_variant} {err:.2e}", flush=True)
        assert err < TOL_FP64
        if equation == "wave":
            continue
        if equation == "euler":
            # the MFMA adjoint kernel (FP64) forms or copies its face geometry
            # the same way; the FP32-storage batch adjoint kernel reads the
            # table: they agree to the FP32 rounding of the latter
            y = np.random.default_rng(order + 40).standard_normal(q.shape)
            qd, yd = (_dev32(a, device) for a in (q, y))
            v64 = plan.rhs_vjp([t.double() for t in qd], [t.double() for t in yd])
            v32 = plan.rhs_vjp(qd, yd)
            ev = max(field_errors(np.stack([t.double().cpu().numpy() for t in v32]),
                                  np.stack([t.cpu().numpy() for t in v64])))
            print(f"    vjp fp64 / fp32-storage: {ev:.2e}", flush=True)
            assert ev < 1e-6
        # the mixed kernels form or copy their face geometry the same way
        q32 = [torch.from_numpy(np.ascontiguousarray(f)).to(device).float() for f in q]
        for precision, tol in (("fp32", 1e-5), ("bf16x3", 1e-4)):
            out = plan.rhs(q32, precision=precision)
            torch.cuda.synchronize()
            gm = np.stack([t.double().cpu().numpy() for t in out])
            em = max(field_errors(gm, ref))
            print(f"    {precision}: {em:.2e}", flush=True)
            assert em < tol


def test_mfma_wave_plan_checks_face_geometry(device):
    """The MFMA wave kernel LIFTs the velocity flux in reference normals,
    which holds when every element's face geometry is the affine image of
    the reference element's (Fscale n_f = rx^T c_f).  A geometry that breaks
    it is refused by an explicit MFMA plan and sent to the LDS-staged kernel
    by an automatic one, which still matches the oracle on it."""
    print("[case] wave P3 MFMA geometry check", flush=True)
    d = make_box_discretization(3, 3)
    geo = d.packed_geometry().copy()
    geo[7, 9 + 4 * 2 + 3] *= 1.5          # one face's Fscale off the identity

    class Bent:
        """The same discretization with the altered face geometry."""
        def __getattr__(self, name):
            return getattr(d, name)

        def packed_geometry(self):
            return geo

    bent = Bent()
    with pytest.raises(ValueError, match="affine image"):
        DGPlan(bent, "wave", 1.0, variant="mfma")
    plan = DGPlan(bent, "wave", 1.0, variant="auto")
    assert plan.kernel_variant == "lds"
    # the mixed wave kernels on the direct flux rely on the same identity:
    # refused on this plan, run on the unaltered one
    q32 = [torch.zeros(d.nelements, d.np, dtype=torch.float32, device=device)
           for _ in range(4)]
    for precision in ("fp32", "bf16x3"):
        with pytest.raises(NotImplementedError, match="reference normals"):
            plan.rhs(q32, precision=precision)
    # the unaltered discretization keeps the MFMA kernel
    assert DGPlan(d, "wave", 1.0, variant="auto").kernel_variant == "mfma"
    # the LDS-staged kernel evaluates the altered operator: compare it with a
    # host evaluation of the same fused forms on the altered geometry
    from oracle import dg_cpu
    q = random_state("wave", d.nodes, seed=3)
    got = _run(plan, q, device)
    ref = dg_cpu.CpuOperator(bent, "wave", 1.0)(q)
    assert rel_l2(got, ref) < TOL_FP64
