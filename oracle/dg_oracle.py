"""CPU FP64 oracle for the DG-FEM right-hand sides (TEST INFRASTRUCTURE ONLY).

This module is the checker for the HIP engine.  Only ``tests/``,
``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py`` may
import it; the engine (``actx_dgfem_suite_amd``) never does.

It restates, in plain numpy and independently of the engine's own setup code,
the operators the reference benchmarks:

* wave: ``wave_operator`` (equations/wave.py:121-200) with the upwind
  ``wave_flux`` (equations/wave.py:80-114) and the mirror boundary state
  ``u+ = -u-, v+ = v-`` (equations/wave.py:152-156);
* Euler: upstream ``grudge.models.euler.EulerOperator(flux_type="lf")`` with
  ``InviscidWallBC`` as wired in equations/euler.py:168-179;
* Maxwell: upstream ``grudge.models.em.MaxwellOperator(epsilon=1, mu=1,
  flux_type=0.5)`` with PEC walls as wired in equations/maxwell.py:252-265.

It follows grudge's *unfused* operator structure -- ``weak_local_div``,
``weak_local_grad``, ``local_grad``, ``face_mass`` and ``inverse_mass`` are
separate steps with explicit mass matrices and Jacobians -- while the engine
uses the reference optimizer's fused forms (inverse mass folded into the
derivative and LIFT matrices, arraycontext/mass_inverse_fuser.py:20-74).

Independent routes used here (so that a shared mistake is unlikely):

* reference operators from a monomial basis with closed-form simplex
  integrals ``int xi^a eta^b zeta^c = a! b! c! / (a+b+c+3)!``, evaluated in
  40-digit arithmetic, instead of the engine's float64 Legendre basis with
  Gauss quadrature;
* metric terms by nodal differentiation of the physical node coordinates
  (grudge's way) instead of vertex differences;
* normals and face areas from cross products of physical face edges, oriented
  away from the element centroid;
* exterior face nodes by nearest-neighbour search on physical coordinates
  (``scipy.spatial.cKDTree``) instead of topological vertex permutations.

Inputs shared with the engine are only the mesh (vertex coordinates and the
element-vertex table) and the reference node set (the warp-and-blend
convention); both are pinned separately by closed-form tests.
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np

NFACES = 4
# local vertex triples of the reference faces (t=-1, s=-1, r+s+t=-1, r=-1)
FACE_VERTICES = ((0, 1, 2), (0, 1, 3), (1, 2, 3), (0, 2, 3))
REF_VERTS = np.array([[-1., -1., -1.], [1., -1., -1.], [-1., 1., -1.], [-1., -1., 1.]])


# {{{ reference operators via monomials

def _monomial_exponents(order, dim):
    out = []
    for tot in range(order + 1):
        if dim == 3:
            for a in range(tot + 1):
                for b in range(tot - a + 1):
                    out.append((a, b, tot - a - b))
        else:
            for a in range(tot + 1):
                out.append((a, tot - a))
    return out


@dataclass(frozen=True)
class OracleRefElement:
    nodes: np.ndarray        # (Np, 3) biunit coordinates
    mass: np.ndarray         # (Np, Np) on the biunit tet
    diff: np.ndarray         # (3, Np, Np) d/dr, d/ds, d/dt
    face_nodes: list         # 4 arrays of volume-node indices
    face_mass_unit: list     # 4 (Nfp, Nfp): int over the unit triangle (area 1/2)


def _mp_matrix(rows):
    import mpmath
    return mpmath.matrix(rows)


def _to_np(m):
    return np.array([[float(m[i, j]) for j in range(m.cols)] for i in range(m.rows)])


_REF_CACHE = {}


def oracle_reference_element(nodes):
    """Monomial-basis nodal operators on the given biunit-tet node set.

    Computed in 40-digit arithmetic (mpmath) from the exact float64 node
    coordinates and closed-form simplex integrals, then rounded to float64, so
    the operators are correct to rounding whatever the monomial basis's
    conditioning.
    """
    import mpmath
    key = nodes.tobytes()
    if key in _REF_CACHE:
        return _REF_CACHE[key]
    mpmath.mp.dps = 40
    npn = nodes.shape[0]
    order = 0
    while (order + 1) * (order + 2) * (order + 3) // 6 < npn:
        order += 1
    exps = _monomial_exponents(order, 3)
    assert len(exps) == npn
    one = mpmath.mpf(1)
    xi = [[(mpmath.mpf(float(c)) + 1) / 2 for c in row] for row in nodes]   # unit simplex
    v = _mp_matrix([[xi[i][0] ** e[0] * xi[i][1] ** e[1] * xi[i][2] ** e[2] for e in exps]
                    for i in range(npn)])
    vinv = v ** -1
    diff = []
    for d in range(3):
        vd = mpmath.matrix(npn, npn)
        for i in range(npn):
            for k, e in enumerate(exps):
                if e[d] == 0:
                    continue
                ee = list(e)
                ee[d] -= 1
                vd[i, k] = e[d] * xi[i][0] ** ee[0] * xi[i][1] ** ee[1] * xi[i][2] ** ee[2]
        diff.append(_to_np(vd * vinv) * 0.5)        # d/dr = (1/2) d/dxi
    def integral(ex):
        num = one
        for a in ex:
            num *= mpmath.factorial(a)
        return num / mpmath.factorial(sum(ex) + len(ex))
    mm = _mp_matrix([[integral(tuple(a + b for a, b in zip(e1, e2))) for e2 in exps]
                     for e1 in exps])
    mass = _to_np(vinv.T * mm * vinv) * 8.0        # biunit volume = 8 * unit volume
    face_nodes, face_mass = [], []
    exps2 = _monomial_exponents(order, 2)
    mm2 = _mp_matrix([[integral(tuple(a + b for a, b in zip(e1, e2))) for e2 in exps2]
                      for e1 in exps2])
    for vs in FACE_VERTICES:
        a0, a1, a2 = REF_VERTS[list(vs)]
        mat = np.stack([a1 - a0, a2 - a0], axis=1)
        uw, *_ = np.linalg.lstsq(mat, (nodes - a0).T, rcond=None)
        resid = np.linalg.norm(mat @ uw - (nodes - a0).T, axis=0)
        idx = np.nonzero(resid < 1e-10)[0]
        face_nodes.append(idx)
        # exact face coordinates: each face node has one of (r, s, t, r+s+t)
        # pinned, so (u, w) follow from two free coordinates of the node
        pts = []
        for i in idx:
            p = [mpmath.mpf(float(c)) for c in nodes[i]]
            # solve p = a0 + u (a1 - a0) + w (a2 - a0) in exact arithmetic
            m2 = mpmath.matrix([[float(a1[k] - a0[k]), float(a2[k] - a0[k])] for k in range(3)])
            rhs = mpmath.matrix([p[k] - float(a0[k]) for k in range(3)])
            sol = mpmath.lu_solve(m2.T * m2, m2.T * rhs)
            pts.append((sol[0], sol[1]))
        v2 = _mp_matrix([[pu ** e[0] * pw ** e[1] for e in exps2] for pu, pw in pts])
        v2inv = v2 ** -1
        face_mass.append(_to_np(v2inv.T * mm2 * v2inv))
    out = OracleRefElement(nodes=nodes, mass=mass, diff=np.stack(diff),
                           face_nodes=face_nodes, face_mass_unit=face_mass)
    _REF_CACHE[key] = out
    return out

# }}}


# {{{ discretization

@dataclass
class OracleDiscretization:
    ref: OracleRefElement
    x: np.ndarray          # (3, K, Np) physical nodes
    rx: np.ndarray         # (K, 3(x), 3(r)) d r / d x
    jac: np.ndarray        # (K,)
    normals: np.ndarray    # (K, 4, 3)
    area: np.ndarray       # (K, 4) physical face areas
    ext_index: np.ndarray  # (K, 4, Nfp) flat exterior volume-node index, -1 on boundary
    nelements: int

    @property
    def np(self):
        return self.ref.nodes.shape[0]

    @property
    def nfp(self):
        return len(self.ref.face_nodes[0])


def build_discretization(vertices, elements, ref_nodes):
    from scipy.spatial import cKDTree
    ref = oracle_reference_element(np.asarray(ref_nodes, dtype=np.float64))
    k = elements.shape[0]
    xv = vertices[elements]                             # (K, 4, 3)
    r, s, t = ref.nodes.T
    lam = np.stack([-(1 + r + s + t) / 2, (1 + r) / 2, (1 + s) / 2, (1 + t) / 2], axis=1)
    x = np.einsum("nv,evx->xen", lam, xv)               # (3, K, Np)
    # metric: dx/dr by nodal differentiation of coordinates (constant per element)
    dxdr = np.einsum("rij,xej->erxi", ref.diff, x).mean(axis=3)  # (K, r, x)
    jm = np.transpose(dxdr, (0, 2, 1))                  # (K, x, r)
    jac = np.linalg.det(jm)
    rx = np.transpose(np.linalg.inv(jm), (0, 2, 1))     # rx[e, x, r] = dr/dx
    centroid = xv.mean(axis=1)
    normals = np.empty((k, NFACES, 3))
    area = np.empty((k, NFACES))
    for f, (i0, i1, i2) in enumerate(FACE_VERTICES):
        cr = np.cross(xv[:, i1] - xv[:, i0], xv[:, i2] - xv[:, i0])
        nrm = np.linalg.norm(cr, axis=1)
        n = cr / nrm[:, None]
        fc = (xv[:, i0] + xv[:, i1] + xv[:, i2]) / 3.0
        sgn = np.sign(np.einsum("ex,ex->e", n, fc - centroid))
        normals[:, f] = n * sgn[:, None]
        area[:, f] = 0.5 * nrm
    npn = ref.nodes.shape[0]
    nfp = len(ref.face_nodes[0])
    fidx = np.stack(ref.face_nodes)                     # (4, Nfp)
    own = np.arange(k)[:, None, None] * npn + fidx[None]  # (K, 4, Nfp)
    xflat = x.reshape(3, -1).T
    fpts = xflat[own]                                   # (K, 4, Nfp, 3)
    # twin faces: a different face with the same centroid
    cent = fpts.mean(axis=2).reshape(-1, 3)
    tree = cKDTree(cent)
    tol = np.min(np.cbrt(np.abs(jac))) * 1e-6
    dist, nn = tree.query(cent, k=2)
    twin = np.full(cent.shape[0], -1, dtype=np.int64)
    for col in range(2):
        ok = (dist[:, col] < tol) & (nn[:, col] != np.arange(cent.shape[0]))
        twin[ok] = nn[ok, col]
    ext = np.full((k * NFACES, nfp), -1, dtype=np.int64)
    has = np.nonzero(twin >= 0)[0]
    a = fpts.reshape(-1, nfp, 3)[has]
    b = fpts.reshape(-1, nfp, 3)[twin[has]]
    dd = np.linalg.norm(a[:, :, None, :] - b[:, None, :, :], axis=3)
    j2 = dd.argmin(axis=2)
    if np.any(np.take_along_axis(dd, j2[:, :, None], axis=2) > tol):
        raise RuntimeError("non-conforming face")
    ext[has] = np.take_along_axis(own.reshape(-1, nfp)[twin[has]], j2, axis=1)
    return OracleDiscretization(ref=ref, x=x, rx=rx, jac=jac, normals=normals,
                                area=area, ext_index=ext.reshape(k, NFACES, nfp),
                                nelements=k)

# }}}


# {{{ grudge-style operators (unfused)

def weak_local_d_dx(d: OracleDiscretization, axis, u):
    """int_E d(phi_i)/dx_axis u dx = J * sum_r (dr/dx) (S_r^T u), S_r = M D_r."""
    m = d.ref.mass
    out = np.zeros_like(u)
    for r in range(3):
        s_t = (m @ d.ref.diff[r]).T
        out += (d.jac * d.rx[:, axis, r])[:, None] * (u @ s_t.T)
    return out


def local_d_dx(d: OracleDiscretization, axis, u):
    out = np.zeros_like(u)
    for r in range(3):
        out += d.rx[:, axis, r][:, None] * (u @ d.ref.diff[r].T)
    return out


def inverse_mass(d: OracleDiscretization, u):
    return (u @ np.linalg.inv(d.ref.mass).T) / d.jac[:, None]


def face_mass(d: OracleDiscretization, flux):
    """sum_f int_{face f} phi_i flux_f dA for flux of shape (K, 4, Nfp)."""
    out = np.zeros((d.nelements, d.np))
    for f in range(NFACES):
        # physical measure = 2*area * (unit-triangle measure)
        contrib = (flux[:, f, :] @ d.ref.face_mass_unit[f].T) * (2.0 * d.area[:, f])[:, None]
        np.add.at(out, (slice(None), d.ref.face_nodes[f]), contrib)
    return out


def interior_trace(d, u):
    fidx = np.stack(d.ref.face_nodes)
    return u[:, fidx]                                   # (K, 4, Nfp)


def exterior_trace(d, u):
    """Neighbour values on interior faces; NaN on boundary faces."""
    flat = u.reshape(-1)
    out = np.full(d.ext_index.shape, np.nan)
    ok = d.ext_index >= 0
    out[ok] = flat[d.ext_index[ok]]
    return out


def boundary_mask(d):
    return d.ext_index < 0


# }}}


# {{{ wave

def wave_rhs(d: OracleDiscretization, u, v, c=1.0):
    """wave_operator of equations/wave.py:121-200 (quad_tag=None)."""
    n = d.normals[:, :, None, :]                        # (K, 4, 1, 3)
    ui, ue = interior_trace(d, u), exterior_trace(d, u)
    vi = np.stack([interior_trace(d, vc) for vc in v], axis=-1)
    ve = np.stack([exterior_trace(d, vc) for vc in v], axis=-1)
    bd = boundary_mask(d)
    # mirror BC: u+ = -u-, v+ = v- (equations/wave.py:152-156)
    ue = np.where(bd, -ui, ue)
    ve = np.where(bd[..., None], vi, ve)
    u_avg, u_diff = 0.5 * (ui + ue), ue - ui
    v_avg, v_diff = 0.5 * (vi + ve), ve - vi
    # wave_flux (equations/wave.py:91-114)
    flux_u = c * (np.sum(v_avg * n, axis=-1) + 0.5 * u_diff)
    v_jump = np.sum(v_diff * n, axis=-1)
    flux_v = c * (u_avg[..., None] * n + 0.5 * v_jump[..., None] * n)
    div_v = sum(weak_local_d_dx(d, i, v[i]) for i in range(3))
    rhs_u = inverse_mass(d, -c * div_v + face_mass(d, flux_u))
    rhs_v = [inverse_mass(d, -c * weak_local_d_dx(d, i, u) + face_mass(d, flux_v[..., i]))
             for i in range(3)]
    return rhs_u, rhs_v

# }}}


# {{{ euler

def euler_volume_flux(q, gamma):
    """Flux tensor F[field][x] of the conserved state (mass, energy, mom[3])."""
    rho, e, mom = q[0], q[1], q[2:5]
    u = mom / rho
    p = (gamma - 1.0) * (e - 0.5 * np.sum(mom * u, axis=0))
    f_mass = [mom[i] for i in range(3)]
    f_energy = [mom[i] * (e + p) / rho for i in range(3)]
    f_mom = [[mom[i] * mom[j] / rho + (p if i == j else 0.0) for j in range(3)]
             for i in range(3)]
    return [f_mass, f_energy] + f_mom


def _euler_wavespeed(q, gamma):
    rho, e, mom = q[0], q[1], q[2:5]
    u = mom / rho
    p = (gamma - 1.0) * (e - 0.5 * np.sum(mom * u, axis=0))
    return np.sqrt(np.sum(u * u, axis=0)) + np.sqrt(gamma * (p / rho))


def euler_numerical_flux(qi, qe, n, gamma):
    """LF flux {F}.n + lam/2 (q- - q+), lam = max of |u|+c over both sides."""
    fi = euler_volume_flux(qi, gamma)
    fe = euler_volume_flux(qe, gamma)
    lam = np.maximum(_euler_wavespeed(qi, gamma), _euler_wavespeed(qe, gamma))
    out = []
    for c in range(5):
        avg_n = sum(0.5 * (fi[c][x] + fe[c][x]) * n[..., x] for x in range(3))
        out.append(avg_n - 0.5 * lam * (qe[c] - qi[c]))
    return np.stack(out)


def euler_rhs(d: OracleDiscretization, q, gamma=1.4):
    """EulerOperator.operator with LF flux and inviscid walls (euler.py:168-179)."""
    q = np.asarray(q)
    flux = euler_volume_flux(q, gamma)
    vol = np.stack([sum(weak_local_d_dx(d, x, flux[c][x]) for x in range(3))
                    for c in range(5)])
    qi = np.stack([interior_trace(d, q[c]) for c in range(5)])
    qe = np.stack([exterior_trace(d, q[c]) for c in range(5)])
    bd = boundary_mask(d)
    n = d.normals                                        # (K, 4, 3)
    nn = np.moveaxis(n, -1, 0)[:, :, :, None]            # (3, K, 4, 1)
    mom_i = qi[2:5]
    mdotn = np.sum(mom_i * nn, axis=0)
    wall = np.concatenate([qi[0:2], mom_i - 2.0 * mdotn[None] * nn])
    qe = np.where(bd[None], wall, qe)
    fstar = euler_numerical_flux(qi, qe, n[:, :, None, :], gamma)
    return np.stack([inverse_mass(d, vol[c] - face_mass(d, fstar[c])) for c in range(5)])

# }}}


# {{{ maxwell

def _cross(a, b):
    return [a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]]


def maxwell_rhs(d: OracleDiscretization, w, alpha=0.5):
    """MaxwellOperator(eps=1, mu=1, flux_type=alpha) with PEC on all walls."""
    w = np.asarray(w)
    e, h = w[0:3], w[3:6]
    grad = [[local_d_dx(d, x, w[c]) for x in range(3)] for c in range(6)]
    curl_e = [grad[2][1] - grad[1][2], grad[0][2] - grad[2][0], grad[1][0] - grad[0][1]]
    curl_h = [grad[5][1] - grad[4][2], grad[3][2] - grad[5][0], grad[4][0] - grad[3][1]]
    wi = np.stack([interior_trace(d, w[c]) for c in range(6)])
    we = np.stack([exterior_trace(d, w[c]) for c in range(6)])
    bd = boundary_mask(d)
    pec = np.concatenate([-wi[0:3], wi[3:6]])
    we = np.where(bd[None], pec, we)
    n = [d.normals[:, :, None, x] for x in range(3)]
    de, dh = we[0:3] - wi[0:3], we[3:6] - wi[3:6]
    n_x_dh = _cross(n, dh)
    n_x_de = _cross(n, de)
    n_x_n_x_de = _cross(n, n_x_de)
    n_x_n_x_dh = _cross(n, n_x_dh)
    flux_e = [-0.5 * n_x_dh[i] + 0.5 * alpha * n_x_n_x_de[i] for i in range(3)]
    flux_h = [0.5 * n_x_de[i] + 0.5 * alpha * n_x_n_x_dh[i] for i in range(3)]
    rhs_e = [curl_h[i] - inverse_mass(d, face_mass(d, flux_e[i])) for i in range(3)]
    rhs_h = [-curl_e[i] - inverse_mass(d, face_mass(d, flux_h[i])) for i in range(3)]
    return np.stack(rhs_e + rhs_h)

# }}}


# {{{ low-storage RK4

# Carpenter & Kennedy (1994), NASA TM-109112, five-stage fourth-order 2N-storage.
LSRK4_A = (0.0,
           -567301805773.0 / 1357537059087.0,
           -2404267990393.0 / 2016746695238.0,
           -3550918686646.0 / 2091501179385.0,
           -1275806237668.0 / 842570457699.0)
LSRK4_B = (1432997174477.0 / 9575080441755.0,
           5161836677717.0 / 13612068292357.0,
           1720146321549.0 / 2090206949498.0,
           3134564353537.0 / 4481467310338.0,
           2277821191437.0 / 14882151754819.0)
LSRK4_C = (0.0,
           1432997174477.0 / 9575080441755.0,
           2526269341429.0 / 6820363183857.0,
           2006345519317.0 / 3224310063776.0,
           2802321613138.0 / 2924317926251.0)


def lsrk4_step(rhs, q, dt, t=0.0):
    """One 5-stage step: res = a_k res + dt f(q); q += b_k res."""
    q = np.array(q, dtype=np.float64, copy=True)
    res = np.zeros_like(q)
    for a, b, c in zip(LSRK4_A, LSRK4_B, LSRK4_C):
        res = a * res + dt * rhs(t + c * dt, q)
        q = q + b * res
    return q

# }}}

# vim: foldmethod=marker
