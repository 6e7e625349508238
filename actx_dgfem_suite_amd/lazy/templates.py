"""What each HIP kernel computes pointwise, for the matcher to compare a
traced operator against (by evaluation at sampled states; never over a
mesh).

A kernel evaluates ``rhs_c = sum_x D(x) W[c][x](q) + sign * LIFT(Fscale *
F*[c](q-, q+, n))`` where ``D(x)`` is the weak derivative ``M^{-1} S_r^T``
(wave, Euler) or the strong ``D_r`` (Maxwell), each contracted with
``dr/dx``.  The templates give ``W`` (per field and axis) and the signed
surface term ``S[c] = sign * F*[c]`` with the boundary state substituted where
``b`` is set:

* wave (equations/wave.py:80-200): ``W[u][x] = -c v_x``, ``W[v_y][x] = -c u
  delta_xy``; ``S = +flux`` with the upwind flux and the mirror state;
* Euler (equations/euler.py:168-174, grudge's ``EulerOperator``):
  ``W[c][x] = F_cx(q)``; ``S = -F*`` (Lax-Friedrichs, walls);
* Maxwell (equations/maxwell.py:57-59, grudge's ``MaxwellOperator``):
  ``W[E_o][x] = sum_a eps_oxa H_a``, ``W[H_o][x] = -sum_a eps_oxa E_a``
  (strong form); ``S = -flux`` (partially upwind, PEC).

The same formulas as the kernels' flux functors and the CPU oracle, restated
over numpy sample arrays; the parameter enters affinely in each probe used
to fit it.
"""

from __future__ import annotations

import numpy as np

# (nfields, uses the weak derivative, surface sign)
SHAPES = {"wave": (4, True, 1.0), "euler": (5, True, -1.0), "maxwell": (6, False, -1.0)}


def _levi(o, x, a):
    if o == x or x == a or o == a:
        return 0
    return 1 if (x - o) % 3 == 1 else -1


def volume(equation: str, q: np.ndarray, p: float) -> np.ndarray:
    """W[c][x] at the sampled states q (nfields, S) -> (nfields, 3, S)."""
    nf = q.shape[0]
    out = np.zeros((nf, 3) + q.shape[1:])
    if equation == "wave":
        for x in range(3):
            out[0, x] = -p * q[1 + x]
            out[1 + x, x] = -p * q[0]
    elif equation == "euler":
        rho, e, m = q[0], q[1], q[2:5]
        u = m / rho
        pr = (p - 1.0) * (e - 0.5 * np.sum(m * u, axis=0))
        for x in range(3):
            out[0, x] = m[x]
            out[1, x] = (e + pr) * u[x]
            for i in range(3):
                out[2 + i, x] = m[i] * u[x] + (pr if i == x else 0.0)
    elif equation == "maxwell":
        for o in range(3):
            for x in range(3):
                for a in range(3):
                    s = _levi(o, x, a)
                    if s:
                        out[o, x] += s * q[3 + a]
                        out[3 + o, x] -= s * q[a]
    else:
        raise KeyError(equation)
    return out


def _boundary_state(equation, qm, n):
    if equation == "wave":
        return np.concatenate([-qm[0:1], qm[1:4]])
    if equation == "euler":
        mn = np.sum(qm[2:5] * n, axis=0)
        return np.concatenate([qm[0:2], qm[2:5] - 2.0 * mn[None] * n])
    return np.concatenate([-qm[0:3], qm[3:6]])


def _cross(a, b):
    return np.stack([a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2],
                     a[0] * b[1] - a[1] * b[0]])


def surface(equation: str, qm, qp, n, b, p: float) -> np.ndarray:
    """S[c] = sign * F*[c](q-, q+ or the boundary state where b, n)."""
    qp = np.where(b[None] > 0.5, _boundary_state(equation, qm, n), qp)
    if equation == "wave":
        u_avg, u_jump = 0.5 * (qm[0] + qp[0]), qp[0] - qm[0]
        v_avg, v_jump = 0.5 * (qm[1:4] + qp[1:4]), qp[1:4] - qm[1:4]
        vj_n = np.sum(v_jump * n, axis=0)
        fu = np.sum(v_avg * n, axis=0) + 0.5 * u_jump
        return p * np.concatenate([fu[None], u_avg[None] * n + 0.5 * vj_n[None] * n])
    if equation == "euler":
        def tensor(q):
            return volume("euler", q, p)

        def speed(q):
            u = q[2:5] / q[0]
            pr = (p - 1.0) * (q[1] - 0.5 * np.sum(q[2:5] * u, axis=0))
            return np.sqrt(np.sum(u * u, axis=0)) + np.sqrt(p * (pr / q[0]))

        fa = 0.5 * (tensor(qm) + tensor(qp))
        lam = np.maximum(speed(qm), speed(qp))
        flux = np.sum(fa * n[None], axis=1) + 0.5 * lam[None] * (qm - qp)
        return -flux
    de, dh = qp[0:3] - qm[0:3], qp[3:6] - qm[3:6]
    n_x_de, n_x_dh = _cross(n, de), _cross(n, dh)
    fe = -0.5 * n_x_dh + 0.5 * p * _cross(n, n_x_de)
    fh = 0.5 * n_x_de + 0.5 * p * _cross(n, n_x_dh)
    return -np.concatenate([fe, fh])


def sample_states(equation: str, nsamp: int, rng) -> np.ndarray:
    """States where every template is defined (Euler: positive density and
    pressure)."""
    nf = SHAPES[equation][0]
    if equation != "euler":
        return rng.standard_normal((nf, nsamp))
    rho = rng.uniform(0.5, 2.0, nsamp)
    u = rng.uniform(-0.5, 0.5, (3, nsamp))
    pr = rng.uniform(0.5, 2.0, nsamp)
    e = pr / 0.4 + 0.5 * rho * np.sum(u * u, axis=0)
    return np.concatenate([rho[None], e[None], rho[None] * u])


# A probe (kind, field, axis) and a state at which each template is affine in
# its parameter with a nonzero slope; the parameter is fitted there and then
# checked everywhere.
def parameter_probe(equation: str):
    if equation == "wave":
        q = np.zeros((4, 1))
        q[1] = 1.0
        return "volume", 0, 0, q          # W[u][x] = -c v_x
    if equation == "euler":
        q = np.array([[1.0], [2.0], [0.0], [0.0], [0.0]])
        return "volume", 2, 0, q          # W[m_x][x] = p = (gamma - 1) E
    q = np.zeros((6, 1))
    return "surface", 1, None, q          # E_y flux with a tangential E jump
