"""GPU parity of the adjoint on triangles (dg_rhs_vjp and dg_rhs_vjp_fp32 on
dim=2 plans, and the LSERK4 reverse sweep over them).

As for tetrahedra, nothing outside this repository computes the adjoint of
the reference's 2D operators (jax is not installed), so it is pinned by what
the mathematics fixes:

* the whole transpose, entry by entry: on a 2^2 x 2 triangle mesh the dense
  operator matrix of the oracle (wave, Maxwell TM) or its central-difference
  Jacobian (Euler) against the matrix the VJP builds from unit cotangents;
* the dot-product identity <L x, y> = <x, L^T y> against the HIP forward
  kernel at every order, on a mesh where blocks run several batches;
* central finite differences of the Euler RHS at P1-P4;
* the FP32 adjoint against the FP64 one, and the dot-product identity of
  whole LSERK4 trajectories with checkpointing.
"""

import numpy as np
import pytest
import torch

from actx_dgfem_suite_amd.plan import DGPlan
from actx_dgfem_suite_amd.timestepping import LSERK4, cfl_timestep, trajectory_vjp
from tests._helpers import discr_pair_2d, nfields, oracle_rhs, random_state

pytestmark = pytest.mark.gpu

PARAM = {"wave": 1.0, "euler": 1.4, "maxwell": 0.5}
SPEED = {"wave": 1.0, "maxwell": 1.0, "euler": 2.0}


def _dev(a, device):
    return [torch.from_numpy(np.ascontiguousarray(f)).to(device) for f in a]


def _np(fields):
    return np.stack([t.cpu().numpy() for t in fields])


def _dot(a, b) -> float:
    return float(sum(torch.sum(x.double() * y.double()) for x, y in zip(a, b)))


def _norm(a) -> float:
    return float(torch.sqrt(sum(torch.sum(x.double() ** 2) for x in a)))


def _shape(equation, d):
    return (nfields(equation, 2), d.nelements, d.np)


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
                                            ("maxwell", 2), ("euler", 1), ("euler", 2)])
def test_2d_vjp_is_the_transpose(equation, order, device):
    print(f"[case] 2d dense transpose {equation} P{order} n=2", flush=True)
    d, od = discr_pair_2d(order, 2)
    shape = _shape(equation, d)
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
def test_2d_dot_product_identity(equation, order, device):
    """n=24: 1,152 triangles, several batches per persistent block at P1."""
    print(f"[case] 2d dot product {equation} P{order} n=24", flush=True)
    d, od = discr_pair_2d(order, 24)
    plan = DGPlan(d, equation, PARAM[equation])
    rng = np.random.default_rng(order)
    shape = _shape(equation, d)
    x = _dev(rng.standard_normal(shape), device)
    y = _dev(rng.standard_normal(shape), device)
    lx = plan.rhs(x)
    lty = plan.rhs_vjp(None, y)
    lhs, rhs = _dot(lx, y), _dot(x, lty)
    rel = abs(lhs - rhs) / (_norm(lx) * _norm(y))
    print(f"  <Lx,y>={lhs:.6e} <x,L^T y>={rhs:.6e} rel={rel:.2e}", flush=True)
    assert rel < 1e-13


def _lf_speed(q, gamma=1.4):
    u2 = sum((q[2 + k] / q[0]) ** 2 for k in range(2))
    p = (gamma - 1.0) * (q[1] - 0.5 * q[0] * u2)
    return np.sqrt(u2) + np.sqrt(gamma * p / q[0])


def _lf_branch_flips(d, q, dq) -> int:
    """Interior face-node pairs whose max(lambda-, lambda+) changes branch
    between q - dq and q + dq (a central difference across that kink
    averages two one-sided derivatives)."""
    em = d.packed_ext_map().reshape(d.nelements, -1)
    own = np.arange(d.nelements)[:, None] * d.np + d.ref.face_nodes.reshape(-1)[None, :]
    a, b = own[em >= 0], em[em >= 0]
    lp, lm = _lf_speed(q + dq).ravel(), _lf_speed(q - dq).ravel()
    return int(np.sum(np.sign(lp[a] - lp[b]) != np.sign(lm[a] - lm[b])))


@pytest.mark.parametrize("order", [1, 2, 3, 4])
def test_2d_euler_vjp_central_difference(order, device):
    """<(L(q+eps v) - L(q-eps v)) / 2 eps, y> = <v, J^T y> for the 2D Euler
    operator, on the first seeded state whose perturbation crosses no
    Lax-Friedrichs wavespeed kink."""
    d, od = discr_pair_2d(order, 6)
    eps = 1e-5
    for seed in range(order, order + 20):
        q = random_state("euler", od.x, seed=seed)
        rng = np.random.default_rng(10 + seed)
        v = rng.standard_normal(q.shape) * np.abs(q).mean(axis=(1, 2), keepdims=True) * 0.1
        if _lf_branch_flips(d, q, eps * v) == 0:
            break
    else:
        pytest.fail("no kink-free state among 20 seeds")
    print(f"[case] 2d euler vjp fd P{order} n=6 seed={seed}", flush=True)
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


@pytest.mark.parametrize("equation", ["wave", "euler", "maxwell"])
@pytest.mark.parametrize("order", [1, 3])
def test_2d_fp32_vjp_against_fp64(equation, order, device):
    """The FP32-storage adjoint computes in FP64 from FP32 inputs: against
    the FP64 adjoint of the same (rounded) inputs it differs by FP32
    rounding of the output only."""
    print(f"[case] 2d fp32 vjp {equation} P{order} n=8", flush=True)
    d, od = discr_pair_2d(order, 8)
    plan = DGPlan(d, equation, PARAM[equation])
    rng = np.random.default_rng(7)
    shape = _shape(equation, d)
    q = random_state(equation, od.x, seed=2).astype(np.float32)
    y = rng.standard_normal(shape).astype(np.float32)
    q32, y32 = _dev(q, device), _dev(y, device)
    q64, y64 = _dev(q.astype(np.float64), device), _dev(y.astype(np.float64), device)
    lin = equation != "euler"
    g32 = plan.rhs_vjp(None if lin else q32, y32)
    g64 = plan.rhs_vjp(None if lin else q64, y64)
    assert all(t.dtype == torch.float32 for t in g32)
    err = _norm([a.double() - b for a, b in zip(g32, g64)]) / _norm(g64)
    print(f"  rel L2 fp32 vs fp64 = {err:.2e}", flush=True)
    assert err < 1e-6

# }}}


# {{{ trajectories

@pytest.mark.parametrize("equation", ["wave", "maxwell"])
def test_2d_trajectory_dot_product_and_checkpointing(equation, device):
    """The LSERK4 step map is linear: <Phi x, y> = <x, Phi^T y>; and the
    gradient does not depend on the checkpoint interval."""
    print(f"[case] 2d trajectory dot product {equation} P3 n=6", flush=True)
    d, od = discr_pair_2d(3, 6)
    plan = DGPlan(d, equation, PARAM[equation])
    st = LSERK4(plan, cfl_timestep(d, SPEED[equation]))
    rng = np.random.default_rng(4)
    shape = _shape(equation, d)
    x = _dev(rng.standard_normal(shape), device)
    y = _dev(rng.standard_normal(shape), device)
    nsteps = 10
    grads = {}
    for every in (1, 4, 10):
        phix, phity = trajectory_vjp(st, x, nsteps, y, checkpoint_every=every)
        grads[every] = phity
        rel = abs(_dot(phix, y) - _dot(x, phity)) / (_norm(phix) * _norm(y))
        print(f"  every={every}: rel={rel:.2e}", flush=True)
        assert rel < 1e-12
    for every in (4, 10):
        for a, b in zip(grads[every], grads[1]):
            assert torch.allclose(a, b, rtol=0, atol=1e-13 * float(b.abs().max()))

# }}}
