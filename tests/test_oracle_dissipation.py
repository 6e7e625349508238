"""Pins the magnitudes of the oracle's dissipative flux terms.

The upwind wave flux carries ``0.5 * u.diff`` and ``0.5 * v_jump * n``
(reference equations/wave.py:98-107), Maxwell's partially upwind flux the
penalty ``flux_type = alpha = 0.5`` (equations/maxwell.py:57-59), and Euler's
Lax-Friedrichs flux ``lambda / 2 * (q- - q+)`` with ``lambda = max(|u| + c)``
over both traces and ``c = sqrt(gamma p / rho)`` (equations/euler.py:168-174).
Polynomial exactness, rest states and conservation all see ``q- = q+`` or are
blind to these coefficients, so these tests fix them by closed forms that
hold for any discrete state:

* wave and Maxwell: with exact quadrature, the energy rate of the DG
  solution equals minus the face jump form.  Summed over both sides of an
  interior face the central terms cancel and the penalty terms leave
  ``-(c/2) int [[u]]^2 + ([[v]].n)^2`` (wave) and
  ``-(alpha/2) int |n x [[E]]|^2 + |n x [[H]]|^2`` (Maxwell); a mirror wall
  (``u+ = -u-``, ``v+ = v-``) leaves ``-c int u-^2`` and a PEC wall
  (``E+ = -E-``, ``H+ = H-``) leaves ``-alpha int |n x E-|^2``;
* Euler: with ``u`` uniform, ``p`` constant and ``rho`` constant per
  element, each element's integrated density rate is
  ``-sum_f |f| ({rho} u.n + lambda_f/2 (rho_e - rho_nbr))`` with
  ``lambda_f = |u| + sqrt(gamma p / min(rho_e, rho_nbr))``, and 0 on walls.

Every integral here is this file's own collapsed Gauss quadrature of the
nodal polynomials, and the geometry comes from the mesh vertices, so neither
the oracle's mass nor its face mass matrices enter the right-hand sides.
Dropping a 1/2, taking lambda from one trace, dropping gamma or |u| from it,
or a wrong wall state each fails one of them (the last test plants those
mistakes and checks that they are caught).
"""

from __future__ import annotations

from itertools import combinations

import numpy as np
import pytest

from oracle import dg_oracle
from tests._helpers import discr_pair, discr_pair_2d, oracle_rhs

NQ = 8   # Gauss points per collapsed direction: exact to degree 13 on P <= 4 products


# {{{ quadrature and interpolation owned by this test

def _gauss01(n):
    g, w = np.polynomial.legendre.leggauss(n)
    return 0.5 * (g + 1.0), 0.5 * w


def _unit_simplex_rule(dim):
    """Points (npts, dim) and weights of the unit simplex {xi >= 0, sum <= 1}
    by the collapsed (Duffy) map of a Gauss tensor rule."""
    a, wa = _gauss01(NQ)
    if dim == 1:
        return a[:, None], wa
    if dim == 2:
        aa, bb = np.meshgrid(a, a, indexing="ij")
        w = np.outer(wa, wa) * (1 - bb)
        pts = np.stack([aa * (1 - bb), bb], axis=-1)
        return pts.reshape(-1, 2), w.ravel()
    aa, bb, cc = np.meshgrid(a, a, a, indexing="ij")
    w = wa[:, None, None] * wa[None, :, None] * wa[None, None, :] * (1 - bb) * (1 - cc) ** 2
    pts = np.stack([aa * (1 - bb) * (1 - cc), bb * (1 - cc), cc], axis=-1)
    return pts.reshape(-1, 3), w.ravel()


def _monomials(pts, order):
    """Monomial Vandermonde (npts, nbasis) of total degree <= order."""
    dim = pts.shape[1]
    cols = []
    for tot in range(order + 1):
        if dim == 1:
            cols.append(pts[:, 0] ** tot)
        elif dim == 2:
            for a in range(tot + 1):
                cols.append(pts[:, 0] ** a * pts[:, 1] ** (tot - a))
        else:
            for a in range(tot + 1):
                for b in range(tot - a + 1):
                    cols.append(pts[:, 0] ** a * pts[:, 1] ** b * pts[:, 2] ** (tot - a - b))
    return np.stack(cols, axis=1)


def _interp_matrix(nodes_param, qpts, order):
    """Values at the quadrature points of the degree-``order`` polynomial
    through the given nodal values: V_q V_nodes^{-1}."""
    return _monomials(qpts, order) @ np.linalg.inv(_monomials(nodes_param, order))


def _order_of(np_, dim):
    n = 0
    while (np.prod([n + 1 + k for k in range(dim)]) // np.prod(range(1, dim + 1))) < np_:
        n += 1
    return n


class Quadrature:
    """Volume and face integrals of nodal fields on an affine simplex mesh.

    ``verts`` (K, dim+1, dim) are the element vertex coordinates; ``x`` (dim,
    K, Np) the physical nodes; ``face_nodes[f]`` the volume-node indices on
    face f.  A face's own parametrisation is found from which element
    vertices lie on it, and its nodes are located in it by least squares."""

    def __init__(self, verts, x, face_nodes):
        self.dim = x.shape[0]
        self.k, self.np = x.shape[1], x.shape[2]
        self.order = _order_of(self.np, self.dim)
        dim = self.dim
        # volume: unit-simplex coordinates of the nodes from element 0's vertices
        v0 = verts[0]
        emat = (v0[1:] - v0[0]).T                        # (dim, dim)
        xi = np.linalg.solve(emat, x[:, 0, :] - v0[0][:, None]).T    # (Np, dim)
        qp, qw = _unit_simplex_rule(dim)
        self.vol_interp = _interp_matrix(xi, qp, self.order)         # (nq, Np)
        self.vol_w = qw
        edges = verts[:, 1:, :] - verts[:, :1, :]                    # (K, dim, dim)
        # physical measure per unit-simplex measure: |det|
        self.vol_scale = np.abs(np.linalg.det(edges))
        # faces
        fq, fw = _unit_simplex_rule(dim - 1)
        self.face_w = fw
        self.face_interp = []      # per (e, f): (nq_face, Nfp)
        self.face_scale = np.empty((self.k, dim + 1))
        self.normals = np.empty((self.k, dim + 1, dim))
        cent = verts.mean(axis=1)
        for e in range(self.k):
            row = []
            for f, fn in enumerate(face_nodes):
                pts = x[:, e, fn].T                                   # (Nfp, dim)
                fv = self._face_vertices(verts[e], pts)
                tang = (fv[1:] - fv[0]).T                             # (dim, dim-1)
                par, *_ = np.linalg.lstsq(tang, (pts - fv[0]).T, rcond=None)
                row.append(_interp_matrix(par.T, fq, self.order))
                if dim == 3:
                    nrm = np.cross(tang[:, 0], tang[:, 1])
                else:
                    nrm = np.array([tang[1, 0], -tang[0, 0]])
                meas = np.linalg.norm(nrm)
                nrm = nrm / meas
                if nrm @ (fv.mean(axis=0) - cent[e]) < 0:
                    nrm = -nrm
                self.normals[e, f] = nrm
                self.face_scale[e, f] = meas           # |cross| or edge length
            self.face_interp.append(row)

    @staticmethod
    def _face_vertices(ev, pts):
        """The dim element vertices lying on the face through ``pts``."""
        best = None
        for combo in combinations(range(len(ev)), len(ev) - 1):
            fv = ev[list(combo)]
            tang = (fv[1:] - fv[0]).T
            par, *_ = np.linalg.lstsq(tang, (pts - fv[0]).T, rcond=None)
            res = np.linalg.norm(tang @ par - (pts - fv[0]).T)
            if best is None or res < best[0]:
                best = (res, fv)
        assert best[0] < 1e-10
        return best[1]

    def volume(self, f, g):
        """sum_e int_E f g dx for nodal fields (K, Np)."""
        fq = f @ self.vol_interp.T
        gq = g @ self.vol_interp.T
        return float(np.sum(self.vol_scale[:, None] * fq * gq * self.vol_w[None, :]))

    def element_integral(self, f):
        """int_E f dx per element, (K,)."""
        return self.vol_scale * ((f @ self.vol_interp.T) @ self.vol_w)

    def face(self, e, f, vals):
        """int over face f of element e of the product of the nodal face
        fields in ``vals`` (each (Nfp,))."""
        out = np.ones(len(self.face_w))
        for v in vals:
            out = out * (self.face_interp[e][f] @ v)
        return float(self.face_scale[e, f] * (out @ self.face_w))

# }}}


def _setup(order, dim):
    if dim == 3:
        d, od = discr_pair(order, 2)
    else:
        d, od = discr_pair_2d(order, 3)
    verts = d.mesh.vertices[d.mesh.elements]
    quad = Quadrature(verts, od.x, od.ref.face_nodes)
    # the quadrature's own normals must agree with the oracle's
    np.testing.assert_allclose(quad.normals, od.normals, atol=1e-13)
    return od, quad


def _traces(od, q):
    """Interior traces (nf, K, F, Nfp), exterior traces (NaN on boundary)."""
    qi = np.stack([dg_oracle.interior_trace(od, c) for c in q])
    qe = np.stack([dg_oracle.exterior_trace(od, c) for c in q])
    return qi, qe


def _state(od, nf, seed):
    rng = np.random.default_rng(seed)
    return rng.standard_normal((nf, od.nelements, od.np))


def _energy_rate(quad, q, rhs):
    return sum(quad.volume(q[c], rhs[c]) for c in range(len(q)))


def _wave_jump_form(od, quad, q, c):
    dim = od.dim
    qi, qe = _traces(od, q)
    bd = od.ext_index < 0
    total = 0.0
    for e in range(od.nelements):
        for f in range(od.nfaces):
            n = quad.normals[e, f]
            ui = qi[0, e, f]
            if bd[e, f, 0]:
                total -= c * quad.face(e, f, [ui, ui])
                continue
            ju = qe[0, e, f] - ui
            jvn = sum((qe[1 + k, e, f] - qi[1 + k, e, f]) * n[k] for k in range(dim))
            # each interior face is visited from both sides: half of -(c/2) per visit
            total -= 0.25 * c * (quad.face(e, f, [ju, ju]) + quad.face(e, f, [jvn, jvn]))
    return total


def _cross3(n, a):
    return [n[1] * a[2] - n[2] * a[1], n[2] * a[0] - n[0] * a[2], n[0] * a[1] - n[1] * a[0]]


def _maxwell_jump_form(od, quad, q, alpha):
    """3D: q = (E, H); 2D TM: q = (E_z, H_x, H_y)."""
    qi, qe = _traces(od, q)
    bd = od.ext_index < 0
    total = 0.0
    for e in range(od.nelements):
        for f in range(od.nfaces):
            nv = quad.normals[e, f]
            if od.dim == 3:
                n = list(nv)
                ei, hi = qi[0:3, e, f], qi[3:6, e, f]
                ee, he = qe[0:3, e, f], qe[3:6, e, f]
            else:
                n = [nv[0], nv[1], 0.0]
                z = np.zeros(od.nfp)
                ei, hi = [z, z, qi[0, e, f]], [qi[1, e, f], qi[2, e, f], z]
                ee, he = [z, z, qe[0, e, f]], [qe[1, e, f], qe[2, e, f], z]
            if bd[e, f, 0]:
                t = _cross3(n, ei)
                total -= alpha * sum(quad.face(e, f, [t[k], t[k]]) for k in range(3))
                continue
            te = _cross3(n, [ee[k] - ei[k] for k in range(3)])
            th = _cross3(n, [he[k] - hi[k] for k in range(3)])
            s = sum(quad.face(e, f, [te[k], te[k]]) + quad.face(e, f, [th[k], th[k]])
                    for k in range(3))
            total -= 0.25 * alpha * s
    return total


@pytest.mark.parametrize("dim,order", [(3, 1), (3, 2), (3, 3), (2, 1), (2, 2), (2, 3), (2, 4)])
def test_wave_energy_rate_is_the_upwind_jump_form(dim, order):
    od, quad = _setup(order, dim)
    c = 1.7
    for seed in range(2):
        q = _state(od, dim + 1, seed)
        rate = _energy_rate(quad, q, oracle_rhs("wave", od, q, c))
        form = _wave_jump_form(od, quad, q, c)
        assert form < 0
        assert abs(rate - form) < 1e-11 * abs(form), (rate, form)


@pytest.mark.parametrize("dim,order", [(3, 1), (3, 2), (3, 3), (2, 1), (2, 2), (2, 3), (2, 4)])
def test_maxwell_energy_rate_is_the_penalty_jump_form(dim, order):
    od, quad = _setup(order, dim)
    nf = 6 if dim == 3 else 3
    for alpha in (0.5, 1.0):
        q = _state(od, nf, 7)
        rate = _energy_rate(quad, q, oracle_rhs("maxwell", od, q, alpha))
        form = _maxwell_jump_form(od, quad, q, alpha)
        assert form < 0
        assert abs(rate - form) < 1e-11 * abs(form), (alpha, rate, form)


def _euler_cellwise_state(od, vel, p, gamma, seed):
    """rho constant per element in [0.5, 2]; u uniform; p constant."""
    rng = np.random.default_rng(seed)
    rho_e = rng.uniform(0.5, 2.0, od.nelements)
    shape = (od.nelements, od.np)
    rho = np.broadcast_to(rho_e[:, None], shape).copy()
    e = p / (gamma - 1.0) + 0.5 * rho * float(np.dot(vel, vel))
    return rho_e, np.stack([rho, e, *(rho * vel[i] for i in range(od.dim))])


def _euler_mass_rate(od, quad, rho_e, vel, p, gamma):
    """-sum_f |f| ({rho} u.n + lambda_f/2 (rho_e - rho_nbr)) per element."""
    out = np.zeros(od.nelements)
    speed = float(np.linalg.norm(vel))
    one = np.ones(od.nfp)
    for e in range(od.nelements):
        for f in range(od.nfaces):
            if od.ext_index[e, f, 0] < 0:
                continue                       # wall: zero mass flux
            nb = od.ext_index[e, f, 0] // od.np
            lam = speed + np.sqrt(gamma * p / min(rho_e[e], rho_e[nb]))
            un = float(quad.normals[e, f] @ vel)
            flux = 0.5 * (rho_e[e] + rho_e[nb]) * un + 0.5 * lam * (rho_e[e] - rho_e[nb])
            out[e] -= flux * quad.face(e, f, [one])
    return out


@pytest.mark.parametrize("dim,order", [(3, 1), (3, 2), (3, 3), (2, 1), (2, 2), (2, 4)])
@pytest.mark.parametrize("moving", [False, True])
def test_euler_lf_wavespeed_closed_form(dim, order, moving):
    """Rest state (u = 0) and uniform flow: per-element integrated density
    rate against the closed form with lambda = |u| + sqrt(gamma p / min rho)."""
    od, quad = _setup(order, dim)
    gamma, p = 1.4, 0.8
    vel = np.array([0.3, -0.2, 0.1][:dim]) if moving else np.zeros(dim)
    rho_e, q = _euler_cellwise_state(od, vel, p, gamma, seed=11)
    rhs = oracle_rhs("euler", od, q, gamma)
    got = quad.element_integral(rhs[0])
    want = _euler_mass_rate(od, quad, rho_e, vel, p, gamma)
    # lambda comes from each side on some faces
    scale = np.abs(want).max()
    assert scale > 0
    np.testing.assert_allclose(got, want, atol=1e-11 * scale)
    if not moving:
        # at rest the only nonzero field is the density
        others = np.abs(rhs[1:]).max()
        assert others < 1e-11 * np.abs(rhs[0]).max()


def test_planted_flux_mistakes_are_caught(monkeypatch):
    """Each plausible mistake in the oracle's dissipative terms moves one of
    the identities above by far more than its tolerance."""
    od, quad = _setup(2, 3)
    c = 1.7
    q = _state(od, 4, 0)
    form = _wave_jump_form(od, quad, q, c)
    # a full jump instead of half a jump in the wave upwind term: the
    # penalty doubles, so the energy rate moves by at least the interior form
    orig_wave = dg_oracle.wave_rhs

    def wave_full_jump(d, u, v, c=1.0):
        ru, rv = orig_wave(d, u, v, c)
        ui, ue = dg_oracle.interior_trace(d, u), dg_oracle.exterior_trace(d, u)
        ue = np.where(d.ext_index < 0, -ui, ue)
        extra = dg_oracle.inverse_mass(d, dg_oracle.face_mass(d, 0.5 * c * (ue - ui)))
        return ru + extra, rv

    ru_bad, rv = wave_full_jump(od, q[0], q[1:4], c)
    rate_bad = _energy_rate(quad, q, np.stack([ru_bad, *rv]))
    assert abs(rate_bad - form) > 0.1 * abs(form)
    # Maxwell: the penalty alpha without its 1/2 is the alpha-doubled operator
    w = _state(od, 6, 7)
    mform = _maxwell_jump_form(od, quad, w, 0.5)
    mrate_bad = _energy_rate(quad, w, oracle_rhs("maxwell", od, w, 1.0))
    assert abs(mrate_bad - mform) > 0.5 * abs(mform)

    # Euler: lambda from the interior trace only, gamma dropped, |u| dropped
    gamma, p = 1.4, 0.8
    vel = np.array([0.3, -0.2, 0.1])
    rho_e, qe = _euler_cellwise_state(od, vel, p, gamma, seed=11)
    want = _euler_mass_rate(od, quad, rho_e, vel, p, gamma)
    tol = 1e-11 * np.abs(want).max()
    orig_speed = dg_oracle._euler_wavespeed
    variants = {
        "one-sided": None,
        "no gamma": lambda qq, g: orig_speed(qq, 1.0),
        "no |u|": lambda qq, g: orig_speed(qq, g) - np.sqrt(
            np.sum((qq[2:5] / qq[0]) ** 2, axis=0)),
    }
    for name, fn in variants.items():
        with monkeypatch.context() as m:
            if name == "one-sided":
                orig_flux = dg_oracle.euler_numerical_flux

                def one_sided(qi, qx, n, g):
                    # lambda of the interior trace alone: max(a, a) = a
                    return orig_flux(qi, qx, n, g) - 0.5 * (
                        np.maximum(orig_speed(qi, g), orig_speed(qx, g))
                        - orig_speed(qi, g)) * (qi - qx)

                m.setattr(dg_oracle, "euler_numerical_flux", one_sided)
            else:
                m.setattr(dg_oracle, "_euler_wavespeed", fn)
            got = quad.element_integral(oracle_rhs("euler", od, qe, gamma)[0])
        assert np.abs(got - want).max() > 1e3 * tol, name

# vim: foldmethod=marker
