"""Maxwell operator: builder mirroring ``get_maxwell_rhs``
(equations/maxwell.py:43-74 of the reference).

The operator is upstream ``grudge.models.em.MaxwellOperator(epsilon=1, mu=1,
flux_type=0.5)`` with PEC walls (equations/maxwell.py:57-59):
``w_t = (curl H, -curl E) - M^{-1} face_mass(flux)`` with the strong-form
local curl and the partially upwind flux of Hesthaven & Warburton p. 433.
"""

from __future__ import annotations

import numpy as np

from actx_dgfem_suite_amd.arraycontext import DGOperator
from actx_dgfem_suite_amd.containers import obj_array
from actx_dgfem_suite_amd.equations._common import box_discretization, make_rhs


def rectangular_cavity_mode(nodes: np.ndarray, t: float = 0.0, mode=(1, 2, 2)):
    """Real part of the TM cavity mode of the unit cube (the state the
    reference takes from grudge's ``get_rectangular_cavity_mode``,
    equations/maxwell.py:61-65)."""
    kx, ky, kz = (m * np.pi for m in mode)
    omega = np.sqrt(kx ** 2 + ky ** 2 + kz ** 2)
    x, y, z = nodes
    sx, cx = np.sin(kx * x), np.cos(kx * x)
    sy, cy = np.sin(ky * y), np.cos(ky * y)
    sz, cz = np.sin(kz * z), np.cos(kz * z)
    tdep = np.exp(-1j * omega * t)
    g2 = kx ** 2 + ky ** 2
    fields = [-kx * kz * cx * sy * sz * tdep / g2,
              -ky * kz * sx * cy * sz * tdep / g2,
              sx * sy * cz * tdep,
              -1j * omega * ky * sx * cy * cz * tdep / g2,
              1j * omega * kx * cx * sy * cz * tdep / g2,
              np.zeros_like(x) * tdep]
    return [np.ascontiguousarray(np.real(f)) for f in fields]


def rectangular_cavity_mode_2d(nodes: np.ndarray, t: float = 0.0, mode=(2, 3)):
    """TM mode of the unit square, fields (E_z, H_x, H_y): E_z = sin(kx x)
    sin(ky y) cos(omega t), H_x = -ky sin(kx x) cos(ky y) sin(omega t) / omega,
    H_y = kx cos(kx x) sin(ky y) sin(omega t) / omega -- an exact solution of
    the 2D operator's PDE with PEC walls (mode (2, 3) as in
    equations/maxwell.py:64-65)."""
    kx, ky = (m * np.pi for m in mode)
    omega = np.hypot(kx, ky)
    x, y = nodes
    return [np.sin(kx * x) * np.sin(ky * y) * np.cos(omega * t),
            -ky * np.sin(kx * x) * np.cos(ky * y) * np.sin(omega * t) / omega,
            kx * np.cos(kx * x) * np.sin(ky * y) * np.sin(omega * t) / omega]


def get_maxwell_rhs(*, actx, dim: int, order: int, ndofs: int | None = None,
                    nel_1d: int | None = None, a=0.0, b=1.0,
                    initial_condition=None, flux_type: float = 0.5):
    """Returns ``(rhs, (t, fields))``; fields are ``[E_x, E_y, E_z, H_x,
    H_y, H_z]`` in 3D and ``[E_z, H_x, H_y]`` (TM) in 2D.  Defaults: unit
    cube with cavity mode (1, 2, 2), unit square with mode (2, 3)."""
    discr = box_discretization(dim, order, ndofs, nel_1d, a, b)
    if initial_condition is None:
        leaves = (rectangular_cavity_mode(discr.nodes) if dim == 3
                  else rectangular_cavity_mode_2d(discr.nodes))
    else:
        leaves = list(initial_condition(discr.nodes))
    fields = actx.from_numpy(obj_array(leaves))
    op = DGOperator("maxwell", discr, float(flux_type), takes_time=True)
    return make_rhs(actx, op), (np.float64(0.5), fields)
