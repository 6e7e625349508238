"""Nodal reference tetrahedron for P1-P4: nodes, derivative, mass, face-mass and
LIFT operators.

This is host-side setup (runs once per operator, not on the hot path).  It
stands in for what the reference obtains from ``modepy`` through
``default_simplex_group_factory(base_dim=3, order=N)``
(equations/wave.py:256-265): a warp-and-blend nodal set on the biunit
tetrahedron with vertices (-1,-1,-1), (1,-1,-1), (-1,1,-1), (-1,-1,1).

Operators are assembled in a total-degree Legendre product basis
``P_a(r) P_b(s) P_c(t)`` (a+b+c <= N), which spans P_N, with mass integrals
taken by a collapsed (Duffy) Gauss-Legendre rule that is exact for the
polynomial degrees involved.  The assembly runs in 40-digit arithmetic
(mpmath) and rounds once to float64, so every operator is correct to
rounding.

Operator meanings (all (Np, Np) unless noted), following the canonical einsum
forms the reference's DG-FEM optimizer produces
(arraycontext/mass_inverse_fuser.py:20-74, transpose_consts_in_einsums.py:5-84):

* ``d_strong[r] = V_r V^{-1}``: strong reference derivative, used by the
  strong-form curl of the Maxwell operator.
* ``d_weak[r] = M^{-1} S_r^T`` with ``S_r = M D_r``: the weak derivative with
  the inverse mass already fused, i.e. the ``ik,rkj->rij`` pre-product of
  mass_inverse_fuser.py:54,72.
* ``lift = M^{-1} E``: (Np, 4*Nfp) face mass fused with the inverse mass, the
  ``ik,kfj->ifj`` pre-product of mass_inverse_fuser.py:35.
"""

from __future__ import annotations

from dataclasses import dataclass
from functools import lru_cache
from math import comb

import numpy as np

NFACES = 4

# Warburton's optimised warp-and-blend blending parameter for 3D, indexed by
# order N (Warburton, J. Eng. Math. 56 (2006) 247-262, Table for tetrahedra).
_ALPHA_OPT_3D = (0.0, 0.0, 0.0, 0.0, 0.1002, 1.1332, 1.5608, 1.3413, 1.2577,
                 1.1603, 1.10153, 0.6080, 0.4523, 0.8856, 0.8717, 0.9655)

# Local vertex triples of each face.  Vertex k of the biunit tet: v0=(-1,-1,-1),
# v1=(1,-1,-1), v2=(-1,1,-1), v3=(-1,-1,1).  Face 0: t=-1, face 1: s=-1,
# face 2: r+s+t=-1, face 3: r=-1.
FACE_VERTICES = ((0, 1, 2), (0, 1, 3), (1, 2, 3), (0, 2, 3))
REF_VERTICES = np.array([[-1., -1., -1.], [1., -1., -1.],
                         [-1., 1., -1.], [-1., -1., 1.]])
# Outward unit normals of the four reference faces.
REF_NORMALS = np.array([[0., 0., -1.], [0., -1., 0.],
                        [1., 1., 1.], [-1., 0., 0.]])
REF_NORMALS[2] /= np.sqrt(3.0)


def np_for_order(order: int) -> int:
    return (order + 1) * (order + 2) * (order + 3) // 6


def nfp_for_order(order: int) -> int:
    return (order + 1) * (order + 2) // 2


# {{{ 1D helpers

def gauss_lobatto_points(n: int) -> np.ndarray:
    """The n+1 Gauss-Lobatto-Legendre points on [-1, 1], ascending."""
    if n == 1:
        return np.array([-1.0, 1.0])
    # interior points are the roots of P_n'
    leg = np.polynomial.legendre.Legendre.basis(n)
    interior = np.sort(np.real(leg.deriv().roots()))
    return np.concatenate([[-1.0], interior, [1.0]])


def _edge_warp(n: int, x: np.ndarray) -> np.ndarray:
    """Warp factor w(x)/(1-x^2) mapping equispaced to GLL points on [-1,1].

    w(x) = sum_i (x_i^GLL - x_i^eq) l_i^eq(x); the division by (1-x^2) is done
    analytically by dropping the two endpoint factors of each Lagrange product
    (the endpoint displacements are zero).
    """
    xeq = np.linspace(-1.0, 1.0, n + 1)
    xgll = gauss_lobatto_points(n)
    out = np.zeros_like(x)
    for i in range(1, n):
        term = np.full_like(x, xgll[i] - xeq[i])
        for j in range(1, n):
            if j != i:
                term = term * (x - xeq[j]) / (xeq[i] - xeq[j])
        # l_i(x)/(1-x^2) = term' * 1/((x_i+1)(1-x_i)) with endpoint factors removed
        term = term / ((xeq[i] + 1.0) * (1.0 - xeq[i]))
        out += term
    return out

# }}}


# {{{ warp & blend nodes

def _equilateral_vertices() -> np.ndarray:
    return np.array([[-1.0, -1.0 / np.sqrt(3.0), -1.0 / np.sqrt(6.0)],
                     [1.0, -1.0 / np.sqrt(3.0), -1.0 / np.sqrt(6.0)],
                     [0.0, 2.0 / np.sqrt(3.0), -1.0 / np.sqrt(6.0)],
                     [0.0, 0.0, 3.0 / np.sqrt(6.0)]])


def _face_shift(n: int, alpha: float, la: np.ndarray, lb: np.ndarray,
                lc: np.ndarray) -> tuple[np.ndarray, np.ndarray]:
    """2D warp on an equilateral triangle with barycentrics (la, lb, lc).

    Each edge contributes its 1D GLL warp, blended by the product of the two
    barycentrics spanning the edge and boosted by (1 + (alpha*l_opp)^2).
    Returns the displacement in the triangle's (tangent1, tangent2) frame.
    """
    w_ab = 4.0 * lb * lc * _edge_warp(n, lc - lb) * (1.0 + (alpha * la) ** 2)
    w_bc = 4.0 * la * lc * _edge_warp(n, la - lc) * (1.0 + (alpha * lb) ** 2)
    w_ca = 4.0 * la * lb * _edge_warp(n, lb - la) * (1.0 + (alpha * lc) ** 2)
    c2, s2 = np.cos(2.0 * np.pi / 3.0), np.sin(2.0 * np.pi / 3.0)
    c4, s4 = np.cos(4.0 * np.pi / 3.0), np.sin(4.0 * np.pi / 3.0)
    dx = w_ab + c2 * w_bc + c4 * w_ca
    dy = s2 * w_bc + s4 * w_ca
    return dx, dy


@lru_cache(maxsize=None)
def warp_blend_nodes(order: int) -> np.ndarray:
    """Warp-and-blend nodes (Warburton 2006) on the biunit tet, shape (Np, 3).

    Nodes are enumerated from the equispaced lattice with r fastest, then s,
    then t.  On edges they coincide with GLL points; the face and interior
    distributions use the optimised blend parameter alpha(N).
    """
    n = order
    if n == 0:
        return np.array([[-0.5, -0.5, -0.5]])
    lat = []
    for k in range(n + 1):
        for j in range(n + 1 - k):
            for i in range(n + 1 - k - j):
                lat.append((-1.0 + 2.0 * i / n, -1.0 + 2.0 * j / n,
                            -1.0 + 2.0 * k / n))
    rst = np.array(lat)
    r, s, t = rst.T
    # barycentrics w.r.t. v3 (t), v2 (s), v0, v1 (r)
    l1 = (1.0 + t) / 2.0
    l2 = (1.0 + s) / 2.0
    l3 = -(1.0 + r + s + t) / 2.0
    l4 = (1.0 + r) / 2.0
    ev = _equilateral_vertices()
    v1, v2, v3, v4 = ev
    tan1 = np.array([v2 - v1, v2 - v1, v3 - v2, v3 - v1])
    tan2 = np.array([v3 - 0.5 * (v1 + v2), v4 - 0.5 * (v1 + v2),
                     v4 - 0.5 * (v2 + v3), v4 - 0.5 * (v1 + v3)])
    tan1 /= np.linalg.norm(tan1, axis=1)[:, None]
    tan2 /= np.linalg.norm(tan2, axis=1)[:, None]

    alpha = _ALPHA_OPT_3D[n] if n < len(_ALPHA_OPT_3D) else 1.0
    tol = 1e-10
    xyz = np.outer(l3, v1) + np.outer(l4, v2) + np.outer(l2, v3) + np.outer(l1, v4)
    shift = np.zeros_like(xyz)
    # (opposite, a, b, c) barycentric roles per equilateral face
    roles = ((l1, l2, l3, l4), (l2, l1, l3, l4), (l3, l1, l4, l2), (l4, l1, l3, l2))
    for face, (la, lb, lc, ld) in enumerate(roles):
        wx, wy = _face_shift(n, alpha, lb, lc, ld)
        blend = lb * lc * ld
        denom = (lb + 0.5 * la) * (lc + 0.5 * la) * (ld + 0.5 * la)
        ok = denom > tol
        blend = np.where(ok, (1.0 + (alpha * la) ** 2) * blend / np.where(ok, denom, 1.0),
                         blend)
        shift += np.outer(blend * wx, tan1[face]) + np.outer(blend * wy, tan2[face])
        on_face = (la < tol) & (((lb > tol).astype(int) + (lc > tol).astype(int)
                                 + (ld > tol).astype(int)) < 3)
        shift[on_face] = (np.outer(wx[on_face], tan1[face])
                          + np.outer(wy[on_face], tan2[face]))
    xyz = xyz + shift
    # back to the biunit tet
    a = 0.5 * np.stack([v2 - v1, v3 - v1, v4 - v1], axis=1)
    b = 0.5 * (v2 + v3 + v4 - v1)
    return np.linalg.solve(a, (xyz - b).T).T

# }}}


# {{{ basis and quadrature (extended precision)

# Operators are assembled in 40-digit arithmetic and rounded once to float64,
# so their accuracy does not depend on the basis' conditioning.
_DPS = 40


def _mp():
    import mpmath
    mpmath.mp.dps = _DPS
    return mpmath


def _total_degree_exponents(order: int, dim: int) -> list[tuple[int, ...]]:
    out = []
    if dim == 3:
        for tot in range(order + 1):
            for a in range(tot, -1, -1):
                for b in range(tot - a, -1, -1):
                    out.append((a, b, tot - a - b))
    else:
        for tot in range(order + 1):
            for a in range(tot, -1, -1):
                out.append((a, tot - a))
    return out


def _legendre_table(order: int, x):
    """[(P_k(x), P_k'(x)) for k = 0..order] for one extended-precision x."""
    mp = _mp()
    p = [mp.mpf(1), x]
    dp = [mp.mpf(0), mp.mpf(1)]
    for n in range(1, order):
        p.append(((2 * n + 1) * x * p[n] - n * p[n - 1]) / (n + 1))
        dp.append(dp[n - 1] + (2 * n + 1) * p[n])
    return list(zip(p[:order + 1], dp[:order + 1]))


def _legendre_product_vandermonde(order: int, pts, with_grad: bool):
    """V[i][k] = prod_d P_{a_kd}(x_id) (and its gradients) as mpmath matrices."""
    mp = _mp()
    dim = len(pts[0])
    exps = _total_degree_exponents(order, dim)
    npts = len(pts)
    v = mp.matrix(npts, len(exps))
    grads = [mp.matrix(npts, len(exps)) for _ in range(dim)] if with_grad else None
    for i, x in enumerate(pts):
        tabs = [_legendre_table(order, x[d]) for d in range(dim)]
        for k, ex in enumerate(exps):
            f = [tabs[d][ex[d]][0] for d in range(dim)]
            prod = mp.mpf(1)
            for d in range(dim):
                prod *= f[d]
            v[i, k] = prod
            if with_grad:
                for d in range(dim):
                    g = tabs[d][ex[d]][1]
                    for d2 in range(dim):
                        if d2 != d:
                            g *= f[d2]
                    grads[d][i, k] = g
    return v, grads


def _gauss_legendre(n: int):
    """n-point Gauss-Legendre rule in extended precision (Newton-refined)."""
    mp = _mp()
    x0, _ = np.polynomial.legendre.leggauss(n)
    xs, ws = [], []
    for guess in x0:
        x = mp.mpf(float(guess))
        for _ in range(8):
            p, dp = _legendre_table(n, x)[n]
            x -= p / dp
        _, dp = _legendre_table(n, x)[n]
        xs.append(x)
        ws.append(2 / ((1 - x * x) * dp * dp))
    return xs, ws


def _collapsed_quadrature(npts: int, dim: int):
    """Gauss-Legendre rule on the biunit simplex via the Duffy collapse.

    Exact for total degree <= 2*npts - 1 - (dim - 1)."""
    x, w = _gauss_legendre(npts)
    pts, wts = [], []
    if dim == 2:
        for i in range(npts):
            for j in range(npts):
                a, b = x[i], x[j]
                pts.append(((1 + a) * (1 - b) / 2 - 1, b))
                wts.append(w[i] * w[j] * (1 - b) / 2)
        return pts, wts
    for i in range(npts):
        for j in range(npts):
            for k in range(npts):
                a, b, c = x[i], x[j], x[k]
                pts.append(((1 + a) * (1 - b) * (1 - c) / 4 - 1,
                            (1 + b) * (1 - c) / 2 - 1, c))
                wts.append(w[i] * w[j] * w[k] * (1 - b) * (1 - c) ** 2 / 8)
    return pts, wts


def _mass_in_basis(order: int, dim: int):
    """Exact Gram matrix int phi_k phi_l of the Legendre-product basis."""
    mp = _mp()
    qp, qw = _collapsed_quadrature(order + 3, dim)
    phi, _ = _legendre_product_vandermonde(order, qp, with_grad=False)
    n = phi.cols
    g = mp.matrix(n, n)
    for k in range(n):
        for l in range(k, n):
            acc = mp.mpf(0)
            for q in range(len(qp)):
                acc += qw[q] * phi[q, k] * phi[q, l]
            g[k, l] = g[l, k] = acc
    return g


def _to_numpy(m) -> np.ndarray:
    return np.array([[float(m[i, j]) for j in range(m.cols)] for i in range(m.rows)])


def _mp_points(a: np.ndarray):
    mp = _mp()
    return [tuple(mp.mpf(float(c)) for c in row) for row in a]

# }}}


@dataclass(frozen=True)
class ReferenceElement:
    order: int
    nodes: np.ndarray          # (Np, 3)
    mass: np.ndarray           # (Np, Np)
    mass_inv: np.ndarray       # (Np, Np)
    d_strong: np.ndarray       # (3, Np, Np)
    d_weak: np.ndarray         # (3, Np, Np)  M^{-1} S_r^T
    face_nodes: np.ndarray     # (4, Nfp) volume node index of each face node
    face_mass: np.ndarray      # (Np, 4*Nfp) E: int_face l_i l^f_j (param. area 2)
    lift: np.ndarray           # (Np, 4*Nfp)

    @property
    def np(self) -> int:
        return self.nodes.shape[0]

    @property
    def nfp(self) -> int:
        return self.face_nodes.shape[1]


def face_param_coords(face: int, rst: np.ndarray) -> np.ndarray:
    """Two coordinates parametrising reference face *face* onto the biunit
    triangle (area 2)."""
    r, s, t = rst.T
    if face == 0:
        return np.stack([r, s], axis=1)
    if face == 1:
        return np.stack([r, t], axis=1)
    if face == 2:
        return np.stack([s, t], axis=1)
    return np.stack([s, t], axis=1)


def _on_face(face: int, rst: np.ndarray, tol: float = 1e-10) -> np.ndarray:
    r, s, t = rst.T
    return np.abs([1.0 + t, 1.0 + s, 1.0 + r + s + t, 1.0 + r][face]) < tol


@lru_cache(maxsize=None)
def reference_element(order: int) -> ReferenceElement:
    if order < 1 or order > 4:
        raise NotImplementedError(f"order {order} (supported: 1-4)")
    nodes = warp_blend_nodes(order)
    npn = nodes.shape[0]
    v, grads = _legendre_product_vandermonde(order, _mp_points(nodes), with_grad=True)
    vinv = v ** -1
    d_strong_mp = [g * vinv for g in grads]
    mass_mp = vinv.T * _mass_in_basis(order, 3) * vinv
    mass_inv_mp = mass_mp ** -1
    d_strong = np.stack([_to_numpy(d) for d in d_strong_mp])
    mass = _to_numpy(mass_mp)
    mass_inv = _to_numpy(mass_inv_mp)

    nfp = nfp_for_order(order)
    face_nodes = np.empty((NFACES, nfp), dtype=np.int32)
    mp = _mp()
    emat_mp = mp.matrix(npn, NFACES * nfp)
    gram2 = _mass_in_basis(order, 2)
    for f in range(NFACES):
        idx = np.nonzero(_on_face(f, nodes))[0]
        if idx.size != nfp:
            raise RuntimeError(f"face {f}: found {idx.size} nodes, expected {nfp}")
        face_nodes[f] = idx
        fcoords = face_param_coords(f, nodes[idx])
        v2, _ = _legendre_product_vandermonde(order, _mp_points(fcoords), with_grad=False)
        v2inv = v2 ** -1
        m2 = v2inv.T * gram2 * v2inv
        for jj, i in enumerate(idx):
            for kk in range(nfp):
                emat_mp[int(i), f * nfp + kk] = m2[jj, kk]
    emat = _to_numpy(emat_mp)
    lift = _to_numpy(mass_inv_mp * emat_mp)
    d_weak = np.stack([_to_numpy(mass_inv_mp * d.T * mass_mp) for d in d_strong_mp])
    return ReferenceElement(order=order, nodes=nodes, mass=mass, mass_inv=mass_inv,
                            d_strong=d_strong, d_weak=d_weak, face_nodes=face_nodes,
                            face_mass=emat, lift=lift)


def n_total_degree(order: int, dim: int) -> int:
    return comb(order + dim, dim)
