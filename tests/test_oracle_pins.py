"""Pins the CPU oracle with closed forms, invariants and brute force.

None of these re-types the oracle's formulas: each checks a property the
mathematics fixes (polynomial exactness of the DG operators, conservation,
upwind dissipation, states at rest, Runge-Kutta order), chosen so that a
dropped term, a wrong sign or a transposed operand fails.
"""

import numpy as np
import pytest

from oracle import dg_oracle
from tests._helpers import discr_pair, interior_elements, oracle_rhs, random_state


def _poly(x, order, seed):
    """Random polynomial of total degree <= order and its gradient."""
    rng = np.random.default_rng(seed)
    exps = [(a, b, c) for a in range(order + 1) for b in range(order + 1 - a)
            for c in range(order + 1 - a - b)]
    coef = rng.standard_normal(len(exps))
    val = np.zeros(x.shape[1:])
    grad = [np.zeros(x.shape[1:]) for _ in range(3)]
    for cf, e in zip(coef, exps):
        val += cf * x[0] ** e[0] * x[1] ** e[1] * x[2] ** e[2]
        for d in range(3):
            if e[d] == 0:
                continue
            ee = list(e)
            ee[d] -= 1
            grad[d] += cf * e[d] * x[0] ** ee[0] * x[1] ** ee[1] * x[2] ** ee[2]
    return val, grad


@pytest.mark.parametrize("order", [1, 2, 3, 4])
def test_wave_polynomial_exactness(order):
    """Continuous polynomial fields: interior RHS = (c div v, c grad u)."""
    _, od = discr_pair(order, 3)
    x = od.x
    c = 1.7
    u, gu = _poly(x, order, 1)
    vs, gv = zip(*[_poly(x, order, 2 + k) for k in range(3)])
    q = np.stack([u, *vs])
    rhs = oracle_rhs("wave", od, q, c)
    inner = interior_elements(od)
    assert inner.size > 0
    div_v = sum(gv[k][k] for k in range(3))
    scale = max(np.abs(div_v).max(), 1.0)
    np.testing.assert_allclose(rhs[0][inner], c * div_v[inner], atol=1e-10 * scale)
    for k in range(3):
        np.testing.assert_allclose(rhs[1 + k][inner], c * gu[k][inner], atol=1e-10 * scale)


@pytest.mark.parametrize("order", [1, 2, 3, 4])
def test_maxwell_polynomial_exactness(order):
    _, od = discr_pair(order, 3)
    x = od.x
    fields, grads = zip(*[_poly(x, order, 10 + k) for k in range(6)])
    q = np.stack(fields)
    rhs = oracle_rhs("maxwell", od, q, 0.5)
    inner = interior_elements(od)

    def curl(g):  # g[c][d] = d field_c / d x_d
        return [g[2][1] - g[1][2], g[0][2] - g[2][0], g[1][0] - g[0][1]]

    curl_e = curl(grads[0:3])
    curl_h = curl(grads[3:6])
    scale = max(max(np.abs(a).max() for a in curl_e + curl_h), 1.0)
    for k in range(3):
        np.testing.assert_allclose(rhs[k][inner], curl_h[k][inner], atol=1e-10 * scale)
        np.testing.assert_allclose(rhs[3 + k][inner], -curl_e[k][inner], atol=1e-10 * scale)


@pytest.mark.parametrize("order", [1, 2, 3])
def test_euler_linear_density_advection(order):
    """rho linear, u and p constant: all fluxes are degree-1 polynomials, so
    interior elements get the exact rho_t = -u.grad rho, m_t = -u (u.grad rho),
    E_t = -|u|^2/2 (u.grad rho)."""
    _, od = discr_pair(order, 3)
    x = od.x
    gamma = 1.4
    grad_rho = np.array([0.1, 0.05, -0.02])
    vel = np.array([0.3, -0.2, 0.1])
    rho = 1.0 + sum(grad_rho[i] * x[i] for i in range(3))
    p = 1.0
    e = p / (gamma - 1) + 0.5 * rho * (vel @ vel)
    q = np.stack([rho, e, *(rho * vel[i] for i in range(3))])
    rhs = oracle_rhs("euler", od, q, gamma)
    inner = interior_elements(od)
    adv = vel @ grad_rho
    np.testing.assert_allclose(rhs[0][inner], -adv, atol=1e-12)
    np.testing.assert_allclose(rhs[1][inner], -0.5 * (vel @ vel) * adv, atol=1e-12)
    for i in range(3):
        np.testing.assert_allclose(rhs[2 + i][inner], -vel[i] * adv, atol=1e-12)


@pytest.mark.parametrize("order", [1, 2, 3, 4])
def test_euler_rest_state_is_steady(order):
    """Constant rho, p and zero velocity with walls: RHS = 0 everywhere."""
    _, od = discr_pair(order, 2)
    shape = od.x.shape[1:]
    q = np.stack([np.full(shape, 1.3), np.full(shape, 2.0 / 0.4), *(np.zeros(shape),) * 3])
    rhs = oracle_rhs("euler", od, q, 1.4)
    # the volume and face terms are each of size ~p / h; they must cancel
    term = np.abs(dg_oracle.inverse_mass(od, dg_oracle.weak_local_d_dx(
        od, 0, np.full(shape, 2.0)))).max()
    assert np.abs(rhs).max() < 1e-13 * term


@pytest.mark.parametrize("order", [1, 2, 3])
def test_euler_conserves_mass_and_energy_with_walls(order):
    _, od = discr_pair(order, 2)
    q = random_state("euler", od.x, seed=3)
    rhs = oracle_rhs("euler", od, q, 1.4)
    w = od.jac[:, None] * (np.ones(od.np) @ od.ref.mass)[None, :]   # integration weights
    for c in (0, 1):
        total = np.sum(w * rhs[c])
        assert abs(total) < 1e-11 * np.sum(w * np.abs(rhs[c])), c


@pytest.mark.parametrize("equation", ["wave", "maxwell"])
@pytest.mark.parametrize("order", [1, 2])
def test_upwind_operator_is_dissipative_brute_force(equation, order):
    """Assemble the dense operator L on a 2^3 x 6 mesh: the M-weighted
    symmetric part of L is negative semi-definite (upwind/PEC/mirror walls),
    and it is strictly dissipative (some negative eigenvalue)."""
    _, od = discr_pair(order, 2)
    nf = 4 if equation == "wave" else 6
    shape = (nf, od.nelements, od.np)
    n = int(np.prod(shape))
    cols = []
    for k in range(n):
        e = np.zeros(n)
        e[k] = 1.0
        cols.append(oracle_rhs(equation, od, e.reshape(shape), 1.0 if equation == "wave" else 0.5).ravel())
    lmat = np.stack(cols, axis=1)
    mblock = np.kron(np.diag(od.jac), od.ref.mass)
    mfull = np.kron(np.eye(nf), mblock)
    sym = mfull @ lmat
    sym = 0.5 * (sym + sym.T)
    ev = np.linalg.eigvalsh(sym)
    scale = np.abs(ev).max()
    assert ev.max() < 1e-10 * scale
    assert ev.min() < -1e-3 * scale


def test_maxwell_central_flux_is_energy_conserving_in_interior():
    """alpha = 0 (central) with PEC walls: the M-weighted operator is skew."""
    _, od = discr_pair(1, 2)
    shape = (6, od.nelements, od.np)
    n = int(np.prod(shape))
    cols = []
    for k in range(n):
        e = np.zeros(n)
        e[k] = 1.0
        cols.append(oracle_rhs("maxwell", od, e.reshape(shape), 0.0).ravel())
    lmat = np.stack(cols, axis=1)
    mfull = np.kron(np.eye(6), np.kron(np.diag(od.jac), od.ref.mass))
    s = mfull @ lmat
    assert np.abs(s + s.T).max() < 1e-10 * np.abs(s).max()


def test_lsrk4_is_fourth_order():
    """Stability polynomial R(z) of the five-stage scheme agrees with e^z to
    O(z^5) and the scheme converges at order 4 on a nonlinear ODE."""
    def amp(z):
        return dg_oracle.lsrk4_step(lambda t, y: z * y, np.array([1.0]), 1.0)[0]

    e1 = abs(amp(0.2) - np.exp(0.2))
    e2 = abs(amp(0.1) - np.exp(0.1))
    assert 25 < e1 / e2 < 40           # ~2^5 local error ratio
    # consistency of stage times: y' = t, exact y = t^2/2 (the published
    # rational coefficients meet the order conditions to ~1e-10)
    y = dg_oracle.lsrk4_step(lambda t, y: np.full_like(y, t), np.array([0.0]), 0.3, t=0.0)
    assert abs(y[0] - 0.045) < 1e-9

    def run(nsteps):
        y = np.array([1.0])
        t = 0.0
        dt = 1.0 / nsteps
        for _ in range(nsteps):
            y = dg_oracle.lsrk4_step(lambda t, y: -y ** 2 + np.cos(t), y, dt, t)
            t += dt
        return y[0]

    ref = run(2000)
    r = (run(10) - ref) / (run(20) - ref)
    assert 13 < r < 19                 # ~2^4 global error ratio
