"""GPU parity of the adjoint (dg_rhs_vjp, dg_lserk_stage_adjoint) and its
autograd bindings.

The reference has no adjoint of its own, and jax (whose ``jax.vjp`` of the
reference's jax:jit RHS would be the natural cross-check) is not installed,
so the adjoint is pinned by what the mathematics fixes:

* the whole transpose, entry by entry: on a 48-tet mesh the dense operator
  matrix of the oracle (linear equations) or its central-difference Jacobian
  (Euler) against the matrix the VJP builds from unit cotangents;
* the dot-product identity <L x, y> = <x, L^T y> against the HIP forward
  kernel, at every order;
* central finite differences of the Euler RHS and of whole LSERK4
  trajectories (objectives quadratic in the final state), and the Maxwell
  adjoint workload at its full size (16 directions, 1e-6).
"""

import ctypes

import numpy as np
import pytest
import torch

from actx_dgfem_suite_amd import autodiff
from actx_dgfem_suite_amd.plan import DGPlan
from actx_dgfem_suite_amd.timestepping import LSERK4, cfl_timestep, reverse_sweep, trajectory_vjp
from actx_dgfem_suite_amd.workloads import WORKLOADS, initial_state, maxwell_target
from tests._helpers import discr_pair, discr_pair_2d, oracle_rhs, random_state

pytestmark = pytest.mark.gpu

PARAM = {"wave": 1.0, "euler": 1.4, "maxwell": 0.5}
NF = {"wave": 4, "euler": 5, "maxwell": 6}
SPEED = {"wave": 1.0, "maxwell": 1.0, "euler": 2.0}


def _dev(a, device):
    return [torch.from_numpy(np.ascontiguousarray(f)).to(device) for f in a]


def _np(fields):
    return np.stack([t.cpu().numpy() for t in fields])


def _dot(a, b) -> float:
    return float(sum(torch.sum(x * y) for x, y in zip(a, b)))


def _norm(a) -> float:
    return float(torch.sqrt(sum(torch.sum(x * x) for x in a)))


# {{{ the whole transpose on a tiny mesh

def _dense_oracle(equation, od, q0, shape):
    """Dense matrix of the oracle RHS (linear) or its central-difference
    Jacobian at q0 (Euler), columns over the flattened state."""
    n = int(np.prod(shape))
    cols = []
    for j in range(n):
        e = np.zeros(n)
        if equation == "euler":
            eps = 1e-6 * max(1.0, abs(q0.ravel()[j]))
            e[j] = eps
            fp = oracle_rhs(equation, od, q0 + e.reshape(shape), PARAM[equation])
            fm = oracle_rhs(equation, od, q0 - e.reshape(shape), PARAM[equation])
            cols.append(((fp - fm) / (2 * eps)).ravel())
        else:
            e[j] = 1.0
            cols.append(oracle_rhs(equation, od, e.reshape(shape), PARAM[equation]).ravel())
    return np.stack(cols, axis=1)


@pytest.mark.parametrize("equation,order", [("wave", 1), ("wave", 2), ("maxwell", 1),
                                            ("maxwell", 2), ("euler", 1)])
def test_vjp_is_the_transpose(equation, order, device):
    print(f"[case] dense transpose {equation} P{order} n=2", flush=True)
    d, od = discr_pair(order, 2)
    nf = NF[equation]
    shape = (nf, d.nelements, d.np)
    q0 = random_state(equation, od.x, seed=3)
    a = _dense_oracle(equation, od, q0, shape)
    plan = DGPlan(d, equation, PARAM[equation])
    qd = _dev(q0, device) if equation == "euler" else None
    n = a.shape[0]
    at = np.empty((n, n))
    lam = plan.empty_fields()
    for i in range(n):
        for t in lam:
            t.zero_()
        c, rest = divmod(i, d.nelements * d.np)
        lam[c].view(-1)[rest] = 1.0
        at[i] = _np(plan.rhs_vjp(qd, lam)).ravel()
    torch.cuda.synchronize()
    scale = np.abs(a).max()
    err = np.abs(at - a).max() / scale
    print(f"  max |VJP matrix - transpose| / max|L| = {err:.2e}", flush=True)
    # Euler: the oracle side is a central difference (truncation ~1e-9)
    assert err < (1e-7 if equation == "euler" else 1e-12)

# }}}


# {{{ dot-product identity against the forward kernel

@pytest.mark.parametrize("equation", ["wave", "maxwell"])
@pytest.mark.parametrize("order", [1, 2, 3, 4])
def test_dot_product_identity(equation, order, device):
    print(f"[case] dot product {equation} P{order} n=4", flush=True)
    d, od = discr_pair(order, 4)
    plan = DGPlan(d, equation, PARAM[equation])
    rng = np.random.default_rng(order)
    shape = (NF[equation], d.nelements, d.np)
    x = _dev(rng.standard_normal(shape), device)
    y = _dev(rng.standard_normal(shape), device)
    lx = plan.rhs(x)
    lty = plan.rhs_vjp(None, y)
    lhs, rhs = _dot(lx, y), _dot(x, lty)
    rel = abs(lhs - rhs) / (_norm(lx) * _norm(y))
    print(f"  <Lx,y>={lhs:.6e} <x,L^T y>={rhs:.6e} rel={rel:.2e}", flush=True)
    assert rel < 1e-13


def _lf_speed(q, gamma=1.4):
    u2 = sum((q[2 + k] / q[0]) ** 2 for k in range(3))
    p = (gamma - 1.0) * (q[1] - 0.5 * q[0] * u2)
    return np.sqrt(u2) + np.sqrt(gamma * p / q[0])


def _lf_branch_flips(d, q, dq) -> int:
    """Interior face-node pairs whose max(lambda-, lambda+) changes branch
    between q - dq and q + dq.  A central difference across that kink
    averages two one-sided derivatives, so it cannot check a VJP there."""
    em = d.packed_ext_map().reshape(d.nelements, -1)
    own = np.arange(d.nelements)[:, None] * d.np + d.ref.face_nodes.reshape(-1)[None, :]
    a, b = own[em >= 0], em[em >= 0]
    lp, lm = _lf_speed(q + dq).ravel(), _lf_speed(q - dq).ravel()
    return int(np.sum(np.sign(lp[a] - lp[b]) != np.sign(lm[a] - lm[b])))


@pytest.mark.parametrize("order", [1, 2, 3, 4])
def test_euler_vjp_central_difference(order, device):
    """<(L(q+eps v) - L(q-eps v)) / 2 eps, y> = <v, J^T y> for the Euler
    operator (truncation O(eps^2)), on the first seeded state whose
    perturbation crosses no Lax-Friedrichs wavespeed kink."""
    d, od = discr_pair(order, 4)
    eps = 1e-5
    for seed in range(order, order + 20):
        q = random_state("euler", od.x, seed=seed)
        rng = np.random.default_rng(10 + seed)
        v = rng.standard_normal(q.shape) * np.abs(q).mean(axis=(1, 2), keepdims=True) * 0.1
        if _lf_branch_flips(d, q, eps * v) == 0:
            break
    else:
        pytest.fail("no kink-free state among 20 seeds")
    print(f"[case] euler vjp fd P{order} n=4 seed={seed}", flush=True)
    plan = DGPlan(d, "euler", PARAM["euler"])
    y = _dev(rng.standard_normal(q.shape), device)
    lp = plan.rhs(_dev(q + eps * v, device))
    lm = plan.rhs(_dev(q - eps * v, device))
    jv = [(a - b) / (2 * eps) for a, b in zip(lp, lm)]
    jty = plan.rhs_vjp(_dev(q, device), y)
    lhs, rhs = _dot(jv, y), _dot(_dev(v, device), jty)
    rel = abs(lhs - rhs) / (_norm(jv) * _norm(y))
    print(f"  <Jv,y>={lhs:.9e} <v,J^T y>={rhs:.9e} rel={rel:.2e}", flush=True)
    assert rel < 1e-8


@pytest.mark.parametrize("equation", ["wave", "maxwell"])
def test_adjoint_space_against_the_oracle_mass(equation, device):
    """dg_adjoint_space: R M^{-1} and R M with M = |J| M_ref per element from
    the oracle's reference mass matrix and Jacobians (R: +1 on u / E, -1 on
    v / H), in place and out of place; the two directions invert each other."""
    print(f"[case] adjoint space {equation} P2 n=3", flush=True)
    d, od = discr_pair(2, 3)
    plan = DGPlan(d, equation, PARAM[equation])
    rng = np.random.default_rng(8)
    x = rng.standard_normal((NF[equation], d.nelements, d.np))
    sign = np.array([1.0] + [-1.0] * 3 if equation == "wave" else [1.0] * 3 + [-1.0] * 3)
    mref, jac = od.ref.mass, np.abs(od.jac)
    want_inv = sign[:, None, None] * np.einsum("ij,cej->cei", np.linalg.inv(mref), x) / jac[None, :, None]
    want_m = sign[:, None, None] * np.einsum("ij,cej->cei", mref, x) * jac[None, :, None]
    got_inv = _np(plan.adjoint_space(_dev(x, device), True))
    xd = _dev(x, device)
    got_m = _np(plan.adjoint_space(xd, False, out=xd))       # in place
    assert np.abs(got_inv - want_inv).max() <= 1e-13 * np.abs(want_inv).max()
    assert np.abs(got_m - want_m).max() <= 1e-13 * np.abs(want_m).max()
    back = _np(plan.adjoint_space(plan.adjoint_space(_dev(x, device), True), False))
    assert np.abs(back - x).max() <= 1e-13 * np.abs(x).max()


@pytest.mark.parametrize("equation", ["wave", "maxwell"])
@pytest.mark.parametrize("order", [2, 4])
def test_linear_vjp_on_both_forward_kernels(equation, order, device):
    """The FP64 linear VJP runs on the plan's forward kernel: the MFMA kernel
    with the mass multiplication fused, or the LDS-staged kernel followed by a
    mass pass.  Both satisfy the dot-product identity against the forward
    RHS and agree with each other to rounding."""
    print(f"[case] linear vjp mfma/lds {equation} P{order} n=3", flush=True)
    d, _ = discr_pair(order, 3)
    rng = np.random.default_rng(9)
    shape = (NF[equation], d.nelements, d.np)
    x = _dev(rng.standard_normal(shape), device)
    y = _dev(rng.standard_normal(shape), device)
    got = {}
    for variant in ("mfma", "lds"):
        plan = DGPlan(d, equation, PARAM[equation], variant=variant)
        assert plan.kernel_variant == variant
        lx, lty = plan.rhs(x), plan.rhs_vjp(None, y)
        rel = abs(_dot(lx, y) - _dot(x, lty)) / (_norm(lx) * _norm(y))
        print(f"  {variant}: rel={rel:.2e}", flush=True)
        assert rel < 1e-13
        got[variant] = lty
    diff = _norm([a - b for a, b in zip(got["mfma"], got["lds"])]) / _norm(got["lds"])
    assert diff < 1e-13, diff


@pytest.mark.parametrize("equation,order", [("wave", 3), ("maxwell", 2), ("euler", 2)])
def test_vjp_in_place_on_the_cotangent(equation, order, device):
    """``out`` may be the cotangent itself, on every VJP path (fused MFMA,
    LDS-staged forward plus mass pass, Euler adjoint kernel): bit-identical
    to the out-of-place result."""
    print(f"[case] vjp in place {equation} P{order} n=3", flush=True)
    d, od = discr_pair(order, 3)
    plan = DGPlan(d, equation, PARAM[equation])
    q = _dev(random_state(equation, od.x, seed=5), device) if equation == "euler" else None
    y = _dev(np.random.default_rng(6).standard_normal((NF[equation], d.nelements, d.np)), device)
    ref = [t.clone() for t in plan.rhs_vjp(q, y)]
    y2 = [t.clone() for t in y]
    got = plan.rhs_vjp(q, y2, out=y2)
    for a, b in zip(got, ref):
        assert torch.equal(a, b)


@pytest.mark.parametrize("equation", ["wave", "maxwell", "euler"])
@pytest.mark.parametrize("order", [1, 2, 3, 4])
def test_vjp_single_partial_tile(equation, order, device):
    """A mesh smaller than one 16-element MFMA tile (6 tets, every face but
    the six internal ones a wall): the fused kernel's padded tile and the
    Euler adjoint kernel's short batch.  Linear: dot-product identity against
    the forward; Euler: central differences of the forward."""
    print(f"[case] vjp {equation} P{order} n=1", flush=True)
    d, od = discr_pair(order, 1)
    plan = DGPlan(d, equation, PARAM[equation])
    rng = np.random.default_rng(order + 20)
    shape = (NF[equation], d.nelements, d.np)
    y = _dev(rng.standard_normal(shape), device)
    if equation != "euler":
        x = _dev(rng.standard_normal(shape), device)
        lx, lty = plan.rhs(x), plan.rhs_vjp(None, y)
        rel = abs(_dot(lx, y) - _dot(x, lty)) / (_norm(lx) * _norm(y))
        print(f"  rel={rel:.2e}", flush=True)
        assert rel < 1e-13
        return
    eps = 1e-5
    for seed in range(order, order + 20):
        q = random_state("euler", od.x, seed=seed)
        v = (np.random.default_rng(30 + seed).standard_normal(q.shape)
             * np.abs(q).mean(axis=(1, 2), keepdims=True) * 0.1)
        if _lf_branch_flips(d, q, eps * v) == 0:
            break
    else:
        pytest.fail("no kink-free state among 20 seeds")
    lp = plan.rhs(_dev(q + eps * v, device))
    lm = plan.rhs(_dev(q - eps * v, device))
    jv = [(a - b) / (2 * eps) for a, b in zip(lp, lm)]
    jty = plan.rhs_vjp(_dev(q, device), y)
    rel = abs(_dot(jv, y) - _dot(_dev(v, device), jty)) / (_norm(jv) * _norm(y))
    print(f"  rel={rel:.2e}", flush=True)
    assert rel < 1e-8


def test_adjoint_space_rejects_euler(device):
    d, _ = discr_pair(1, 2)
    plan = DGPlan(d, "euler", PARAM["euler"])
    with pytest.raises(NotImplementedError, match="not linear"):
        plan.adjoint_space(plan.empty_fields(), True)


def test_vjp_many_batches_and_workspace_reuse(device):
    """A mesh where every persistent block runs several batches; two calls in
    a row with different cotangents reuse the workspace."""
    print("[case] dot product wave P3 n=12 (reuse)", flush=True)
    d, od = discr_pair(3, 12)
    plan = DGPlan(d, "wave", 1.0)
    rng = np.random.default_rng(5)
    shape = (4, d.nelements, d.np)
    x = _dev(rng.standard_normal(shape), device)
    lx = plan.rhs(x)
    for seed in (1, 2):
        y = _dev(np.random.default_rng(seed).standard_normal(shape), device)
        rel = abs(_dot(lx, y) - _dot(x, plan.rhs_vjp(None, y))) / (_norm(lx) * _norm(y))
        assert rel < 1e-13, (seed, rel)

# }}}


# {{{ autograd

@pytest.mark.parametrize("equation", ["wave", "euler", "maxwell"])
def test_autograd_rhs(equation, device):
    print(f"[case] autograd rhs {equation} P2 n=3", flush=True)
    d, od = discr_pair(2, 3)
    plan = DGPlan(d, equation, PARAM[equation])
    q0 = random_state(equation, od.x, seed=1)
    q = [t.requires_grad_() for t in _dev(q0, device)]
    out = autodiff.rhs(plan, q)
    for a, b in zip(out, plan.rhs(_dev(q0, device))):
        assert torch.equal(a.detach(), b)
    rng = np.random.default_rng(2)
    w = _dev(rng.standard_normal((NF[equation], d.nelements, d.np)), device)
    loss = sum(torch.sum(a * b) for a, b in zip(out, w))
    grads = torch.autograd.grad(loss, q)
    ref = plan.rhs_vjp(_dev(q0, device) if equation == "euler" else None, w)
    for g, r in zip(grads, ref):
        assert torch.equal(g, r)

# }}}


# {{{ trajectories

def _objective(q, target):
    return 0.5 * sum(torch.sum((a - b) ** 2) for a, b in zip(q, target))


@pytest.mark.parametrize("equation", ["wave", "maxwell"])
def test_trajectory_dot_product_and_checkpointing(equation, device):
    """The step map is linear: <Phi x, y> = <x, Phi^T y>; and the gradient
    does not depend on the checkpoint interval."""
    print(f"[case] trajectory dot product {equation} P2 n=3", flush=True)
    d, od = discr_pair(2, 3)
    plan = DGPlan(d, equation, PARAM[equation])
    st = LSERK4(plan, cfl_timestep(d, SPEED[equation]))
    rng = np.random.default_rng(4)
    shape = (NF[equation], d.nelements, d.np)
    x = _dev(rng.standard_normal(shape), device)
    y = _dev(rng.standard_normal(shape), device)
    nsteps = 12
    grads = {}
    for every in (1, 5, 12):
        phix, phity = trajectory_vjp(st, x, nsteps, y, checkpoint_every=every)
        grads[every] = phity
        rel = abs(_dot(phix, y) - _dot(x, phity)) / (_norm(phix) * _norm(y))
        print(f"  every={every}: rel={rel:.2e}", flush=True)
        assert rel < 1e-12
    for every in (5, 12):
        for a, b in zip(grads[every], grads[1]):
            assert torch.allclose(a, b, rtol=0, atol=1e-13 * float(b.abs().max()))


def test_euler_trajectory_central_difference(device):
    """Gradient of J = 1/2 ||q(T) - q_tgt||^2 after 8 Euler LSERK4 steps
    against central differences in 4 directions.  The perturbed trajectories
    are checked to cross no Lax-Friedrichs wavespeed kink at any stage (a
    continuous state would tie max(lambda-, lambda+) at every face)."""
    print("[case] euler trajectory fd P2 n=3", flush=True)
    d, od = discr_pair(2, 3)
    plan = DGPlan(d, "euler", PARAM["euler"])
    st = LSERK4(plan, cfl_timestep(d, SPEED["euler"]))
    nsteps, eps = 8, 1e-6
    q0 = random_state("euler", od.x, seed=6)
    target = _dev(random_state("euler", od.x, seed=7), device)
    q = [t.requires_grad_() for t in _dev(q0, device)]
    qt = autodiff.trajectory(st, q, nsteps, checkpoint_every=3)
    grads = torch.autograd.grad(_objective(qt, target), q)
    rng = np.random.default_rng(8)

    def stage_states(qs):
        out = []
        for _ in range(nsteps):
            st.step(qs, out)
        return qs, out

    for k in range(4):
        v = rng.standard_normal(q0.shape) * np.abs(q0).mean(axis=(1, 2), keepdims=True) * 0.05
        qp, sp = stage_states(_dev(q0 + eps * v, device))
        qm, sm = stage_states(_dev(q0 - eps * v, device))
        flips = sum(_lf_branch_flips(d, 0.5 * (_np(a) + _np(b)), 0.5 * (_np(a) - _np(b)))
                    for a, b in zip(sp, sm))
        assert flips == 0, f"direction {k} crosses {flips} wavespeed kinks"
        fd = float(_objective(qp, target) - _objective(qm, target)) / (2 * eps)
        ad = _dot(grads, _dev(v, device))
        print(f"  dir {k}: fd={fd:.10e} adjoint={ad:.10e} rel={abs(fd - ad) / abs(ad):.2e}",
              flush=True)
        assert abs(fd - ad) <= 1e-6 * abs(ad)


def _run(st, q, nsteps):
    st.run(q, nsteps)
    return q


def test_maxwell_adjoint_workload_gradient(device):
    """The Maxwell adjoint workload at its full size: 25^3 x 6 tets, P4,
    100 LSERK4 steps at the CFL step, checkpoints every 10, objective
    J = 1/2 ||E(T) - E_tgt||^2; the gradient w.r.t. the initial state
    against central differences on 16 random directions (seed 2)."""
    w = WORKLOADS["maxwell_p4"]
    print(f"[case] maxwell adjoint workload K={6 * w.nel_1d ** 3} P4 100 steps", flush=True)
    d = w.discretization()
    plan = DGPlan(d, "maxwell", w.param)
    st = LSERK4(plan, cfl_timestep(d, 1.0))
    q0 = initial_state(w, d)
    e_tgt = _dev(maxwell_target(d)[:3], device)

    def objective(q):
        return 0.5 * sum(torch.sum((a - b) ** 2) for a, b in zip(q[:3], e_tgt))

    q = [t.requires_grad_() for t in _dev(q0, device)]
    qt = autodiff.trajectory(st, q, 100, checkpoint_every=10)
    j0 = objective(qt)
    grads = torch.autograd.grad(j0, q)
    print(f"  J={float(j0.detach()):.10e}", flush=True)
    rng = np.random.default_rng(2)
    worst = 0.0
    for k in range(16):
        v = rng.standard_normal(q0.shape)
        eps = 1e-3
        jp = objective(_run(st, _dev(q0 + eps * v, device), 100))
        jm = objective(_run(st, _dev(q0 - eps * v, device), 100))
        fd = float(jp - jm) / (2 * eps)
        ad = _dot(grads, _dev(v, device))
        rel = abs(fd - ad) / abs(ad)
        worst = max(worst, rel)
        print(f"  dir {k:2d}: fd={fd:.12e} adjoint={ad:.12e} rel={rel:.2e}", flush=True)
    assert worst < 1e-6

# }}}


@pytest.mark.parametrize("equation", ["wave", "maxwell"])
def test_upwind_energy_does_not_grow(equation, device):
    """Upwind fluxes dissipate: the discrete energy sum_e J_e q^T M q stays
    below its initial value along an LSERK4 trajectory at the CFL step."""
    print(f"[case] energy {equation} P3 n=4 60 steps", flush=True)
    d, od = discr_pair(3, 4)
    plan = DGPlan(d, equation, 1.0)
    st = LSERK4(plan, cfl_timestep(d, 1.0))
    q = _dev(random_state(equation, od.x, seed=9), device)
    mass = torch.from_numpy(d.ref.mass).to(device)
    jac = torch.from_numpy(np.ascontiguousarray(d.geometry.jac)).to(device)

    def energy():
        return float(sum(torch.sum(jac[:, None] * (f @ mass) * f) for f in q))

    e0 = energy()
    for _ in range(60):
        st.step(q)
        cur = energy()
        # one RK step of a dissipative but non-normal operator may not be
        # strictly contractive; the energy never exceeds its initial value
        assert cur <= e0 * (1 + 1e-12), (e0, cur)
    assert cur < e0


@pytest.mark.parametrize("precision", ["fp32", "bf16x3"])
def test_fp32_adjoint_matches_fp64(precision, device):
    """The FP32 adjoint (dg_rhs_vjp_fp32, dg_lserk_stage_adjoint_fp32) on the
    wave P4 adjoint workload's set-up at a reduced mesh (10^3 x 6 tets):
    forward in the mixed form, 40 LSERK4 steps, gradient of the final-time
    misfit of u against the FP64 forward + FP64 adjoint."""
    from actx_dgfem_suite_amd.timestepping import misfit_gradient
    w = WORKLOADS["wave_p4_large"]
    print(f"[case] fp32 adjoint wave P4 n=10 forward {precision}", flush=True)
    d = w.discretization(nel_1d=10)
    plan = DGPlan(d, "wave", 1.0)
    dt = cfl_timestep(d, 1.0)
    q0 = _dev(initial_state(w, d), device)
    target = _dev(initial_state(w, d, seed=1)[:1], device)
    j64, g64 = misfit_gradient(LSERK4(plan, dt), q0, target, 40, 20)
    j32, g32 = misfit_gradient(LSERK4(plan, dt, precision=precision), q0, target, 40, 20,
                               adjoint_dtype=torch.float32)
    assert all(t.dtype == torch.float32 for t in g32)
    err = float(torch.linalg.norm(g32[0].double() - g64[0]) / torch.linalg.norm(g64[0]))
    print(f"  J fp64={j64:.10e} mixed={j32:.10e}; grad_u rel L2 = {err:.2e}", flush=True)
    assert abs(j32 - j64) <= 1e-4 * j64
    assert err < 1e-4
    # single VJP: FP32 against FP64 on the same cotangent
    y = _dev(np.random.default_rng(3).standard_normal((4, d.nelements, d.np)), device)
    v64 = plan.rhs_vjp(None, y)
    v32 = plan.rhs_vjp(None, [t.float() for t in y])
    e1 = max(float(torch.linalg.norm(a.double() - b) / torch.linalg.norm(b))
             for a, b in zip(v32, v64))
    assert e1 < 1e-6, e1


@pytest.mark.parametrize("dim", [3, 2])
@pytest.mark.parametrize("equation", ["wave", "maxwell"])
def test_fp32_adjoint_space_sweep_matches_fp64(dim, equation, device):
    """Float32 cotangents of the linear operators sweep back in the variables
    of plan.adjoint_space on the FP32 mixed RHS: the FP32 transform against
    the FP64 one (both directions), and a 10-step reverse sweep against the
    FP64 sweep of the same cotangent."""
    print(f"[case] fp32 adjoint-space sweep {equation} dim={dim} P3", flush=True)
    d, _ = discr_pair(3, 4) if dim == 3 else discr_pair_2d(3, 6)
    plan = DGPlan(d, equation, PARAM[equation])
    st = LSERK4(plan, cfl_timestep(d, SPEED[equation]))
    nf = plan.nfields
    y = _dev(np.random.default_rng(5).standard_normal((nf, d.nelements, d.np)), device)
    y32 = [t.float() for t in y]
    for to_adjoint in (True, False):
        a64 = plan.adjoint_space(y, to_adjoint)
        a32 = plan.adjoint_space(y32, to_adjoint)
        assert all(t.dtype == torch.float32 for t in a32)
        e = max(float(torch.linalg.norm(a.double() - b) / torch.linalg.norm(b))
                for a, b in zip(a32, a64))
        print(f"  transform to_adjoint={to_adjoint}: rel err {e:.2e}", flush=True)
        assert e < 1e-6
    g64 = reverse_sweep(st, {}, 10, y)
    g32 = reverse_sweep(st, {}, 10, y32)
    assert all(t.dtype == torch.float32 for t in g32)
    e = max(float(torch.linalg.norm(a.double() - b) / torch.linalg.norm(b)) for a, b in zip(g32, g64))
    print(f"  10-step sweep: rel err {e:.2e}", flush=True)
    assert e < 1e-5


@pytest.mark.parametrize("equation", ["wave", "maxwell"])
@pytest.mark.parametrize("order", [1, 2, 3, 4])
def test_fused_reverse_stage_matches_unfused(equation, order, device):
    """dg_rhs_lserk_stage_adjoint (the reverse stage update applied in the
    MFMA kernel's output phase) against dg_rhs followed by
    dg_lserk_stage_adjoint: two reverse steps in the adjoint variables on a
    mesh with a partial last tile."""
    print(f"[case] fused reverse stage {equation} P{order} n=3", flush=True)
    d, _ = discr_pair(order, 3)
    plan = DGPlan(d, equation, PARAM[equation], variant="mfma")
    fused, plain = LSERK4(plan, 1e-3, fused=True), LSERK4(plan, 1e-3, fused=False)
    assert fused._fused_adjoint and not plain._fused_adjoint
    y = _dev(np.random.default_rng(order).standard_normal((plan.nfields, d.nelements, d.np)),
             device)
    a, b = [t.clone() for t in y], [t.clone() for t in y]
    for _ in range(2):
        fused.step_vjp_adjoint_space(a)
        plain.step_vjp_adjoint_space(b)
    e = max(float(torch.linalg.norm(x - z) / torch.linalg.norm(z)) for x, z in zip(a, b))
    print(f"  rel err fused vs unfused {e:.2e}", flush=True)
    assert e < 1e-14
    s = plan.empty_fields()
    with pytest.raises(ValueError, match="s_out must not overlap s_in"):
        plan.rhs_lserk_stage_adjoint(0.5, 0.5, 1e-3, s, s, a)
    with pytest.raises(ValueError, match="q_hat must not overlap s_in"):
        plan.rhs_lserk_stage_adjoint(0.5, 0.5, 1e-3, s, plan.empty_fields(), s)
    if order == 1:
        de, _ = discr_pair(1, 2)
        eu = DGPlan(de, "euler", PARAM["euler"], variant="mfma")
        f = eu.empty_fields()
        with pytest.raises(NotImplementedError, match="not linear"):
            eu.rhs_lserk_stage_adjoint(0.5, 0.5, 1e-3, f, eu.empty_fields(), eu.empty_fields())


def test_fp32_euler_vjp_matches_fp64(device):
    d, od = discr_pair(2, 4)
    plan = DGPlan(d, "euler", PARAM["euler"])
    q = random_state("euler", od.x, seed=2)
    y = np.random.default_rng(4).standard_normal(q.shape)
    v64 = plan.rhs_vjp(_dev(q, device), _dev(y, device))
    v32 = plan.rhs_vjp([t.float() for t in _dev(q, device)], [t.float() for t in _dev(y, device)])
    e = max(float(torch.linalg.norm(a.double() - b) / torch.linalg.norm(b)) for a, b in zip(v32, v64))
    print(f"[case] fp32 euler vjp: rel err {e:.2e}", flush=True)
    assert e < 1e-5


@pytest.mark.parametrize("equation", ["wave", "euler"])
@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_vjp_rejects_aliasing(equation, precision, device):
    """dg_rhs_vjp / dg_rhs_vjp_fp32 reject, before launching anything, a
    q_bar that shares memory with the workspace (or, for Euler, with q) and a
    workspace that overlaps q (Euler) or another field of rhs_bar; a
    workspace equal to rhs_bar field by field is accepted (the mass pass is
    safe in place) and gives the plan's own result."""
    print(f"[case] vjp aliasing {equation} {precision}", flush=True)
    from actx_dgfem_suite_amd import _lib
    d, od = discr_pair(2, 2)
    plan = DGPlan(d, equation, PARAM[equation])
    dt = torch.float64 if precision == "fp64" else torch.float32
    fn = plan.lib.dg_rhs_vjp if precision == "fp64" else plan.lib.dg_rhs_vjp_fp32
    nf = NF[equation]
    rng = np.random.default_rng(3)
    q = [t.to(dt) for t in _dev(random_state(equation, od.x, seed=1), device)]
    y = [t.to(dt) for t in _dev(rng.standard_normal((nf, d.nelements, d.np)), device)]
    qb = plan.empty_fields(dt)
    ws = plan.empty_fields(dt)
    stream = ctypes.c_void_p(torch.cuda.current_stream(device).cuda_stream)

    def call(qq, yy, out, wk):
        p = _lib.pointer_array
        return fn(plan._handle, p([t.data_ptr() for t in qq]), p([t.data_ptr() for t in yy]),
                  p([t.data_ptr() for t in out]), p([t.data_ptr() for t in wk]), stream)

    def rejected(qq, yy, out, wk, match):
        rc = call(qq, yy, out, wk)
        assert rc == 1, rc
        assert match in plan.lib.dg_last_error_string().decode()

    rejected(q, y, ws, ws, "q_bar must not overlap workspace")
    rejected(q, y, qb, [ws[1]] + ws[1:], "] and [1] overlap")
    rejected(q, y, qb, [y[1]] + ws[1:], "workspace[0] overlaps rhs_bar[1]")
    if equation == "euler":
        rejected(q, y, q, ws, "q_bar must not overlap q")
        rejected(q, y, qb, q, "workspace must not overlap q")
    torch.cuda.synchronize()
    ref = [t.clone() for t in plan.rhs_vjp(q if equation == "euler" else None, y)]
    y2 = [t.clone() for t in y]
    assert call(q, y2, qb, y2) == 0
    torch.cuda.synchronize()
    for a, b in zip(qb, ref):
        assert torch.equal(a, b)
