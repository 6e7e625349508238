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
  flux_type=0.5)`` with PEC walls as wired in equations/maxwell.py:57-59.

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

Triangles (2D) take the same routes: ``int xi^a eta^b = a! b! / (a+b+2)!``
over the unit triangle, edge mass matrices from the exact interval moments
``1/(a+b+1)``, outward normals by rotating edge vectors away from the element
centroid, twin edges by coordinates.  The operators are written over
``range(dim)``; 2D Maxwell is the TM restriction (E_z, H_x, H_y) of the 3D
operator.

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
    nodes: np.ndarray        # (Np, dim) biunit coordinates
    mass: np.ndarray         # (Np, Np) on the biunit simplex
    diff: np.ndarray         # (dim, Np, Np) d/dr, d/ds(, d/dt)
    face_nodes: list         # dim+1 arrays of volume-node indices
    face_mass_unit: list     # (Nfp, Nfp) per face: int over the unit triangle
                             # (area 1/2) in 3D, the unit interval in 2D


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


# local vertex pairs of the reference triangle's edges (s=-1, r+s=0, r=-1)
EDGE_VERTICES = ((0, 1), (1, 2), (2, 0))
REF_VERTS_2D = np.array([[-1., -1.], [1., -1.], [-1., 1.]])


def oracle_reference_triangle(nodes):
    """Monomial-basis nodal operators on a biunit-triangle node set, as
    :func:`oracle_reference_element` does for the tetrahedron: closed-form
    integrals over the unit triangle (int xi^a eta^b = a! b! / (a+b+2)!) and
    the unit interval for the edges, in 40-digit arithmetic.  ``diff`` has 2
    rows; ``face_mass_unit`` integrates over each edge parametrised on [0, 1]."""
    import mpmath
    key = ("tri", nodes.tobytes())
    if key in _REF_CACHE:
        return _REF_CACHE[key]
    mpmath.mp.dps = 40
    npn = nodes.shape[0]
    order = 0
    while (order + 1) * (order + 2) // 2 < npn:
        order += 1
    exps = _monomial_exponents(order, 2)
    assert len(exps) == npn
    xi = [[(mpmath.mpf(float(c)) + 1) / 2 for c in row] for row in nodes]
    v = _mp_matrix([[xi[i][0] ** e[0] * xi[i][1] ** e[1] for e in exps] for i in range(npn)])
    vinv = v ** -1
    diff = []
    for d in range(2):
        vd = mpmath.matrix(npn, npn)
        for i in range(npn):
            for k, e in enumerate(exps):
                if e[d] == 0:
                    continue
                ee = list(e)
                ee[d] -= 1
                vd[i, k] = e[d] * xi[i][0] ** ee[0] * xi[i][1] ** ee[1]
        diff.append(_to_np(vd * vinv) * 0.5)
    mm = _mp_matrix([[mpmath.factorial(e1[0] + e2[0]) * mpmath.factorial(e1[1] + e2[1])
                      / mpmath.factorial(e1[0] + e2[0] + e1[1] + e2[1] + 2)
                      for e2 in exps] for e1 in exps])
    mass = _to_np(vinv.T * mm * vinv) * 4.0        # biunit area = 4 * unit area
    face_nodes, face_mass = [], []
    nfp = order + 1
    m1 = _mp_matrix([[mpmath.mpf(1) / (a + b + 1) for b in range(nfp)] for a in range(nfp)])
    for a_, b_ in EDGE_VERTICES:
        a0, a1 = REF_VERTS_2D[a_], REF_VERTS_2D[b_]
        tang = a1 - a0
        rel = nodes - a0
        cross = rel[:, 0] * tang[1] - rel[:, 1] * tang[0]
        idx = np.nonzero(np.abs(cross) < 1e-10)[0]
        face_nodes.append(idx)
        # edge coordinate in [0, 1], exact from the node coordinates
        k = 0 if abs(tang[0]) > 0 else 1
        ts = [(mpmath.mpf(float(nodes[i][k])) - float(a0[k])) / float(tang[k]) for i in idx]
        v1 = _mp_matrix([[t ** e for e in range(nfp)] for t in ts])
        v1inv = v1 ** -1
        face_mass.append(_to_np(v1inv.T * m1 * v1inv))
    out = OracleRefElement(nodes=nodes, mass=mass, diff=np.stack(diff),
                           face_nodes=face_nodes, face_mass_unit=face_mass)
    _REF_CACHE[key] = out
    return out

# }}}


# {{{ discretization

@dataclass
class OracleDiscretization:
    ref: OracleRefElement
    x: np.ndarray          # (dim, K, Np) physical nodes
    rx: np.ndarray         # (K, dim(x), dim(r)) d r / d x
    jac: np.ndarray        # (K,)
    normals: np.ndarray    # (K, nfaces, dim)
    area: np.ndarray       # (K, nfaces) physical face areas (edge lengths in 2D)
    ext_index: np.ndarray  # (K, nfaces, Nfp) flat exterior volume-node index, -1 on boundary
    nelements: int
    # physical measure of a face per unit of the reference face measure the
    # face_mass_unit matrices use (3D: unit triangle of area 1/2; 2D: [0, 1])
    face_measure: np.ndarray = None

    @property
    def dim(self):
        return self.x.shape[0]

    @property
    def nfaces(self):
        return self.dim + 1

    @property
    def np(self):
        return self.ref.nodes.shape[0]

    @property
    def nfp(self):
        return len(self.ref.face_nodes[0])


def _twin_faces(own, xflat, tol):
    """Exterior node index of every face node by matching faces with equal
    centroids and then nodes with equal coordinates (k-d tree)."""
    from scipy.spatial import cKDTree
    nf_total, nfp = own.reshape(-1, own.shape[-1]).shape
    fpts = xflat[own.reshape(-1, nfp)]                  # (K*F, Nfp, dim)
    cent = fpts.mean(axis=1)
    tree = cKDTree(cent)
    dist, nn = tree.query(cent, k=2)
    twin = np.full(cent.shape[0], -1, dtype=np.int64)
    for col in range(2):
        ok = (dist[:, col] < tol) & (nn[:, col] != np.arange(cent.shape[0]))
        twin[ok] = nn[ok, col]
    ext = np.full((nf_total, nfp), -1, dtype=np.int64)
    has = np.nonzero(twin >= 0)[0]
    a = fpts[has]
    b = fpts[twin[has]]
    dd = np.linalg.norm(a[:, :, None, :] - b[:, None, :, :], axis=3)
    j2 = dd.argmin(axis=2)
    if np.any(np.take_along_axis(dd, j2[:, :, None], axis=2) > tol):
        raise RuntimeError("non-conforming face")
    ext[has] = np.take_along_axis(own.reshape(-1, nfp)[twin[has]], j2, axis=1)
    return ext.reshape(own.shape)


def build_discretization_2d(vertices, elements, ref_nodes):
    """Triangles: metric by nodal differentiation, normals by rotating edge
    vectors away from the centroid, twin edges by coordinates."""
    ref = oracle_reference_triangle(np.asarray(ref_nodes, dtype=np.float64))
    k = elements.shape[0]
    xv = vertices[elements]                             # (K, 3, 2)
    r, s = ref.nodes.T
    lam = np.stack([-(r + s) / 2, (1 + r) / 2, (1 + s) / 2], axis=1)
    x = np.einsum("nv,evx->xen", lam, xv)               # (2, K, Np)
    dxdr = np.einsum("rij,xej->erxi", ref.diff, x).mean(axis=3)   # (K, r, x)
    jm = np.transpose(dxdr, (0, 2, 1))
    jac = np.linalg.det(jm)
    rx = np.transpose(np.linalg.inv(jm), (0, 2, 1))
    centroid = xv.mean(axis=1)
    normals = np.empty((k, 3, 2))
    length = np.empty((k, 3))
    for f, (i0, i1) in enumerate(EDGE_VERTICES):
        t = xv[:, i1] - xv[:, i0]
        ln = np.linalg.norm(t, axis=1)
        n = np.stack([t[:, 1], -t[:, 0]], axis=1) / ln[:, None]
        mid = 0.5 * (xv[:, i0] + xv[:, i1])
        sgn = np.sign(np.einsum("ex,ex->e", n, mid - centroid))
        normals[:, f] = n * sgn[:, None]
        length[:, f] = ln
    npn = ref.nodes.shape[0]
    fidx = np.stack(ref.face_nodes)                     # (3, Nfp)
    own = np.arange(k)[:, None, None] * npn + fidx[None]
    tol = np.min(np.sqrt(np.abs(jac))) * 1e-6
    ext = _twin_faces(own, x.reshape(2, -1).T, tol)
    return OracleDiscretization(ref=ref, x=x, rx=rx, jac=jac, normals=normals,
                                area=length, ext_index=ext, nelements=k,
                                face_measure=length)


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
                                nelements=k, face_measure=2.0 * area)

# }}}


# {{{ grudge-style operators (unfused)

def weak_local_d_dx(d: OracleDiscretization, axis, u):
    """int_E d(phi_i)/dx_axis u dx = J * sum_r (dr/dx) (S_r^T u), S_r = M D_r."""
    m = d.ref.mass
    out = np.zeros_like(u)
    for r in range(d.dim):
        s_t = (m @ d.ref.diff[r]).T
        out += (d.jac * d.rx[:, axis, r])[:, None] * (u @ s_t.T)
    return out


def local_d_dx(d: OracleDiscretization, axis, u):
    out = np.zeros_like(u)
    for r in range(d.dim):
        out += d.rx[:, axis, r][:, None] * (u @ d.ref.diff[r].T)
    return out


def inverse_mass(d: OracleDiscretization, u):
    return (u @ np.linalg.inv(d.ref.mass).T) / d.jac[:, None]


def face_mass(d: OracleDiscretization, flux):
    """sum_f int_{face f} phi_i flux_f dA for flux of shape (K, nfaces, Nfp)."""
    out = np.zeros((d.nelements, d.np))
    for f in range(d.nfaces):
        # physical measure = face_measure * (unit reference-face measure)
        contrib = (flux[:, f, :] @ d.ref.face_mass_unit[f].T) * d.face_measure[:, f][:, None]
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
    """wave_operator of equations/wave.py:121-200 (quad_tag=None), in 2D or 3D."""
    n = d.normals[:, :, None, :]                        # (K, F, 1, dim)
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
    div_v = sum(weak_local_d_dx(d, i, v[i]) for i in range(d.dim))
    rhs_u = inverse_mass(d, -c * div_v + face_mass(d, flux_u))
    rhs_v = [inverse_mass(d, -c * weak_local_d_dx(d, i, u) + face_mass(d, flux_v[..., i]))
             for i in range(d.dim)]
    return rhs_u, rhs_v

# }}}


# {{{ euler

def euler_volume_flux(q, gamma):
    """Flux tensor F[field][x] of the conserved state (mass, energy, mom[dim])."""
    dim = len(q) - 2
    rho, e, mom = q[0], q[1], q[2:2 + dim]
    u = mom / rho
    p = (gamma - 1.0) * (e - 0.5 * np.sum(mom * u, axis=0))
    f_mass = [mom[i] for i in range(dim)]
    f_energy = [mom[i] * (e + p) / rho for i in range(dim)]
    f_mom = [[mom[i] * mom[j] / rho + (p if i == j else 0.0) for j in range(dim)]
             for i in range(dim)]
    return [f_mass, f_energy] + f_mom


def _euler_wavespeed(q, gamma):
    dim = len(q) - 2
    rho, e, mom = q[0], q[1], q[2:2 + dim]
    u = mom / rho
    p = (gamma - 1.0) * (e - 0.5 * np.sum(mom * u, axis=0))
    return np.sqrt(np.sum(u * u, axis=0)) + np.sqrt(gamma * (p / rho))


def euler_numerical_flux(qi, qe, n, gamma):
    """LF flux {F}.n + lam/2 (q- - q+), lam = max of |u|+c over both sides."""
    dim = len(qi) - 2
    fi = euler_volume_flux(qi, gamma)
    fe = euler_volume_flux(qe, gamma)
    lam = np.maximum(_euler_wavespeed(qi, gamma), _euler_wavespeed(qe, gamma))
    out = []
    for c in range(dim + 2):
        avg_n = sum(0.5 * (fi[c][x] + fe[c][x]) * n[..., x] for x in range(dim))
        out.append(avg_n - 0.5 * lam * (qe[c] - qi[c]))
    return np.stack(out)


def euler_rhs(d: OracleDiscretization, q, gamma=1.4):
    """EulerOperator.operator with LF flux and inviscid walls (euler.py:168-179),
    in 2D or 3D."""
    q = np.asarray(q)
    dim = d.dim
    nfield = dim + 2
    flux = euler_volume_flux(q, gamma)
    vol = np.stack([sum(weak_local_d_dx(d, x, flux[c][x]) for x in range(dim))
                    for c in range(nfield)])
    qi = np.stack([interior_trace(d, q[c]) for c in range(nfield)])
    qe = np.stack([exterior_trace(d, q[c]) for c in range(nfield)])
    bd = boundary_mask(d)
    n = d.normals                                        # (K, F, dim)
    nn = np.moveaxis(n, -1, 0)[:, :, :, None]            # (dim, K, F, 1)
    mom_i = qi[2:2 + dim]
    mdotn = np.sum(mom_i * nn, axis=0)
    wall = np.concatenate([qi[0:2], mom_i - 2.0 * mdotn[None] * nn])
    qe = np.where(bd[None], wall, qe)
    fstar = euler_numerical_flux(qi, qe, n[:, :, None, :], gamma)
    return np.stack([inverse_mass(d, vol[c] - face_mass(d, fstar[c])) for c in range(nfield)])

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


def maxwell_rhs_2d(d: OracleDiscretization, w, alpha=0.5):
    """2D MaxwellOperator (eps = mu = 1, flux_type alpha, PEC) in TM form: the
    fields (E_z, H_x, H_y) of the 3D operator with E_x = E_y = H_z = 0 and
    no z-dependence; the 3D flux formulas with n_z = 0."""
    w = np.asarray(w)
    ez, hx, hy = w
    dez = [local_d_dx(d, x, ez) for x in range(2)]
    dhx_dy = local_d_dx(d, 1, hx)
    dhy_dx = local_d_dx(d, 0, hy)
    wi = np.stack([interior_trace(d, w[c]) for c in range(3)])
    we = np.stack([exterior_trace(d, w[c]) for c in range(3)])
    bd = boundary_mask(d)
    pec = np.concatenate([-wi[0:1], wi[1:3]])
    we = np.where(bd[None], pec, we)
    zero = np.zeros_like(wi[0])
    n = [d.normals[:, :, None, 0], d.normals[:, :, None, 1], zero]
    de = [zero, zero, we[0] - wi[0]]
    dh = [we[1] - wi[1], we[2] - wi[2], zero]
    n_x_dh = _cross(n, dh)
    n_x_de = _cross(n, de)
    n_x_n_x_de = _cross(n, n_x_de)
    n_x_n_x_dh = _cross(n, n_x_dh)
    flux_ez = -0.5 * n_x_dh[2] + 0.5 * alpha * n_x_n_x_de[2]
    flux_hx = 0.5 * n_x_de[0] + 0.5 * alpha * n_x_n_x_dh[0]
    flux_hy = 0.5 * n_x_de[1] + 0.5 * alpha * n_x_n_x_dh[1]
    rhs_ez = dhy_dx - dhx_dy - inverse_mass(d, face_mass(d, flux_ez))
    rhs_hx = -dez[1] - inverse_mass(d, face_mass(d, flux_hx))
    rhs_hy = dez[0] - inverse_mass(d, face_mass(d, flux_hy))
    return np.stack([rhs_ez, rhs_hx, rhs_hy])

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
