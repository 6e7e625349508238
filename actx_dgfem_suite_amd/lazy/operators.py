"""The three operators written against the grudge-shaped layer, in the
structure of the reference's code, so that tracing them gives the DAG a
lazy array context would receive from the reference's ``rhs`` closures:

* ``wave_operator``: equations/wave.py:121-200 with ``wave_flux``
  (:80-114): ``inverse_mass(-c weak_div v, -c weak_grad u) + face_mass(
  flux on the boundary face set + flux on the interior face set)``, the
  boundary pair being ``(w, (-u, v))``;
* ``euler_operator``: grudge's ``EulerOperator`` with the LF flux and
  ``InviscidWallBC``, as configured at equations/euler.py:168-174:
  ``inverse_mass(weak_div F(q) - face_mass(F*))``;
* ``maxwell_operator``: grudge's ``MaxwellOperator(epsilon=1, mu=1,
  flux_type=alpha)`` with PEC walls (equations/maxwell.py:57-59): the strong
  curl minus ``inverse_mass(face_mass(flux))``.

``get_{wave,euler,maxwell}_rhs`` mirror the reference builders'
signatures and return ``(rhs, args)`` whose ``rhs`` closes over the
operator, for :class:`HIPDGFEMLazyArrayContext` to trace.
"""

from __future__ import annotations

import numpy as np

from actx_dgfem_suite_amd.containers import (
    ConservedEulerField,
    WaveState,
    flatten_fields,
    obj_array,
)
from actx_dgfem_suite_amd.lazy import grudge_ops as op
from actx_dgfem_suite_amd.lazy.grudge_ops import LazyDiscretizationCollection, TracePair


def _faces(dcoll, bdry_flux, int_flux):
    """Both face sets' fluxes on all faces (grudge's projections to all_faces)."""
    return [op.project_boundary_to_all_faces(dcoll, b) + op.project_interior_to_all_faces(dcoll, i)
            for b, i in zip(bdry_flux, int_flux)]


def _dot(a, b):
    return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]


# {{{ wave

def wave_flux(dcoll, c, u: TracePair, v: list) -> list:
    """equations/wave.py:80-114: (n.{v} + [[u]]/2, {u} n + ([[v]].n) n / 2), times c."""
    n = op.normal(dcoll)
    v_avg = [p.avg for p in v]
    v_jump = _dot([p.diff for p in v], n)
    flux_u = _dot(v_avg, n) + 0.5 * u.diff
    flux_v = [u.avg * n[x] + 0.5 * v_jump * n[x] for x in range(3)]
    return [c * f for f in [flux_u, *flux_v]]


def wave_operator(actx, dcoll, c: float, w: WaveState) -> WaveState:
    u, v = w.u, list(w.v)
    dir_u = op.boundary_trace(dcoll, u)
    dir_v = [op.boundary_trace(dcoll, vx) for vx in v]
    bdry = wave_flux(dcoll, c, TracePair(int=dir_u, ext=-dir_u),
                     [TracePair(int=d, ext=d) for d in dir_v])
    inner = wave_flux(dcoll, c, op.interior_trace_pair(dcoll, u),
                      [op.interior_trace_pair(dcoll, vx) for vx in v])
    flux = _faces(dcoll, bdry, inner)
    vol = [-c * op.weak_local_div(dcoll, v), *(-c * g for g in op.weak_local_grad(dcoll, u))]
    out = [op.inverse_mass(dcoll, a + op.face_mass(dcoll, f)) for a, f in zip(vol, flux)]
    return WaveState(u=out[0], v=obj_array(out[1:]))

# }}}


# {{{ euler

def euler_flux_tensor(actx, q: list, gamma: float) -> list:
    """F[c][x] of the conserved state (rho, E, rho u)."""
    rho, e, m = q[0], q[1], q[2:5]
    u = [mi / rho for mi in m]
    p = (gamma - 1.0) * (e - 0.5 * _dot(m, u))
    return ([list(m), [(e + p) * ux for ux in u]]
            + [[m[i] * u[x] + (p if i == x else 0.0) for x in range(3)] for i in range(3)])


def _wavespeed(actx, q, gamma):
    rho, e, m = q[0], q[1], q[2:5]
    u = [mi / rho for mi in m]
    p = (gamma - 1.0) * (e - 0.5 * _dot(m, u))
    return actx.np.sqrt(_dot(u, u)) + actx.np.sqrt(gamma * (p / rho))


def euler_lf_flux(actx, dcoll, qi: list, qe: list, gamma: float) -> list:
    """{F}.n + lam/2 (q- - q+), lam = max(|u| + c) over both traces."""
    n = op.normal(dcoll)
    fi = euler_flux_tensor(actx, qi, gamma)
    fe = euler_flux_tensor(actx, qe, gamma)
    lam = actx.np.maximum(_wavespeed(actx, qi, gamma), _wavespeed(actx, qe, gamma))
    return [_dot([0.5 * (fi[c][x] + fe[c][x]) for x in range(3)], n)
            + 0.5 * lam * (qi[c] - qe[c]) for c in range(5)]


def euler_operator(actx, dcoll, gamma: float, q: ConservedEulerField) -> ConservedEulerField:
    fields = flatten_fields(q)
    flux = euler_flux_tensor(actx, fields, gamma)
    n = op.normal(dcoll)
    # walls: the exterior state reflects the normal momentum
    qb = [op.boundary_trace(dcoll, f) for f in fields]
    mn = _dot(qb[2:5], n)
    wall = qb[0:2] + [qb[2 + x] - 2.0 * mn * n[x] for x in range(3)]
    bdry = euler_lf_flux(actx, dcoll, qb, wall, gamma)
    pairs = [op.interior_trace_pair(dcoll, f) for f in fields]
    inner = euler_lf_flux(actx, dcoll, [p.int for p in pairs], [p.ext for p in pairs], gamma)
    fstar = _faces(dcoll, bdry, inner)
    out = [op.inverse_mass(dcoll, op.weak_local_div(dcoll, flux[c]) - op.face_mass(dcoll, fstar[c]))
           for c in range(5)]
    return ConservedEulerField(mass=out[0], energy=out[1], momentum=obj_array(out[2:5]))

# }}}


# {{{ maxwell

def _cross(a, b):
    return [a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]]


def maxwell_flux(dcoll, alpha, e: list, h: list) -> list:
    """Partially upwind flux with jumps [[.]] = ext - int:
    E: -n x [[H]] / 2 + alpha n x (n x [[E]]) / 2,
    H:  n x [[E]] / 2 + alpha n x (n x [[H]]) / 2."""
    n = list(op.normal(dcoll))
    de = [p.diff for p in e]
    dh = [p.diff for p in h]
    n_x_de, n_x_dh = _cross(n, de), _cross(n, dh)
    n_x_n_x_de, n_x_n_x_dh = _cross(n, n_x_de), _cross(n, n_x_dh)
    return ([-0.5 * n_x_dh[i] + 0.5 * alpha * n_x_n_x_de[i] for i in range(3)]
            + [0.5 * n_x_de[i] + 0.5 * alpha * n_x_n_x_dh[i] for i in range(3)])


def maxwell_operator(actx, dcoll, alpha: float, w) -> np.ndarray:
    fields = flatten_fields(w)
    grads = [op.local_grad(dcoll, f) for f in fields]          # grads[c][x]

    def curl(g):
        return [g[2][1] - g[1][2], g[0][2] - g[2][0], g[1][0] - g[0][1]]

    curl_e, curl_h = curl(grads[0:3]), curl(grads[3:6])
    eb = [op.boundary_trace(dcoll, f) for f in fields[0:3]]
    hb = [op.boundary_trace(dcoll, f) for f in fields[3:6]]
    # PEC: E+ = -E-, H+ = H-
    bdry = maxwell_flux(dcoll, alpha, [TracePair(int=x, ext=-x) for x in eb],
                        [TracePair(int=x, ext=x) for x in hb])
    pairs = [op.interior_trace_pair(dcoll, f) for f in fields]
    inner = maxwell_flux(dcoll, alpha, pairs[0:3], pairs[3:6])
    flux = _faces(dcoll, bdry, inner)
    strong = curl_h + [-x for x in curl_e]
    return obj_array([strong[c] - op.inverse_mass(dcoll, op.face_mass(dcoll, flux[c]))
                      for c in range(6)])

# }}}


# {{{ builders (signatures of the reference's get_*_rhs)

def _discretization(order, ndofs, nel_1d, a, b):
    from actx_dgfem_suite_amd.equations._common import box_discretization
    return box_discretization(3, order, ndofs, nel_1d, a, b)


def get_wave_rhs(*, actx, dim: int = 3, order: int, ndofs: int | None = None,
                 nel_1d: int | None = None, a=-0.5, b=0.5, initial_condition=None,
                 c: float = 1.0):
    """``(rhs, (fields,))`` with ``rhs(w) = wave_operator(actx, dcoll, c, w)``
    (equations/wave.py:239-285); the default state is the ``bump`` IC."""
    if dim != 3:
        raise NotImplementedError("the lazy operators are written for tetrahedra")
    from actx_dgfem_suite_amd.equations.wave import bump
    discr = _discretization(order, ndofs, nel_1d, a, b)
    dcoll = LazyDiscretizationCollection(actx, discr)
    if initial_condition is None:
        u = bump(discr.nodes)
        vs = [np.zeros_like(u) for _ in range(3)]
    else:
        u, *vs = initial_condition(discr.nodes)
    fields = actx.from_numpy(WaveState(u=u, v=obj_array(vs)))

    def rhs(w):
        return wave_operator(actx, dcoll, c, w)

    return rhs, (fields,)


def get_euler_rhs(*, actx, dim: int = 3, order: int, ndofs: int | None = None,
                  nel_1d: int | None = None, a=-0.5, b=0.5, initial_condition=None,
                  gamma: float = 1.4):
    """``(rhs, (t, fields))`` with ``rhs(t, q) = euler_operator(...)``
    (equations/euler.py:140-186)."""
    if dim != 3:
        raise NotImplementedError("the lazy operators are written for tetrahedra")
    from actx_dgfem_suite_amd.equations.euler import acoustic_pulse_condition
    discr = _discretization(order, ndofs, nel_1d, a, b)
    dcoll = LazyDiscretizationCollection(actx, discr)
    if initial_condition is None:
        mass, energy, mom = acoustic_pulse_condition(discr.nodes, gamma)
    else:
        mass, energy, *mom = initial_condition(discr.nodes)
    fields = actx.from_numpy(ConservedEulerField(mass=mass, energy=energy,
                                                 momentum=obj_array(mom)))

    def rhs(t, q):
        return euler_operator(actx, dcoll, gamma, q)

    return rhs, (np.float64(0.5), fields)


def get_maxwell_rhs(*, actx, dim: int = 3, order: int, ndofs: int | None = None,
                    nel_1d: int | None = None, a=0.0, b=1.0, initial_condition=None,
                    flux_type: float = 0.5):
    """``(rhs, (t, fields))`` with ``rhs(t, w) = maxwell_operator(...)``
    (equations/maxwell.py:43-74); the default state is the (1, 2, 2) cavity
    mode."""
    if dim != 3:
        raise NotImplementedError("the lazy operators are written for tetrahedra")
    from actx_dgfem_suite_amd.equations.maxwell import rectangular_cavity_mode
    discr = _discretization(order, ndofs, nel_1d, a, b)
    dcoll = LazyDiscretizationCollection(actx, discr)
    leaves = (rectangular_cavity_mode(discr.nodes) if initial_condition is None
              else list(initial_condition(discr.nodes)))
    fields = actx.from_numpy(obj_array(leaves))

    def rhs(t, w):
        return maxwell_operator(actx, dcoll, flux_type, w)

    return rhs, (np.float64(0.5), fields)

# }}}
