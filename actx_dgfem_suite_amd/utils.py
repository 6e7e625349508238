"""Mesh sizing helpers mirroring utils.py:46-74 of the reference."""

from __future__ import annotations

from math import ceil, sqrt

_NP_3D = {1: 4, 2: 10, 3: 20, 4: 35}


def _ceil_cbrt(v: float) -> int:
    """Exact ceil(cbrt(v)) for v >= 0."""
    n = max(int(round(v ** (1.0 / 3.0))), 0)
    while n ** 3 < v:
        n += 1
    while n > 0 and (n - 1) ** 3 >= v:
        n -= 1
    return n
_NP_2D = {1: 3, 2: 6, 3: 10, 4: 15}


def get_nel_1d_for_regular_rect_mesh(dim: int, order: int, ndofs: int) -> int:
    """Hexahedra per axis so that 6*n^3 tets (3D) hold at least ``ndofs``
    nodes per field (utils.py:15-43)."""
    if dim == 3:
        if order not in _NP_3D:
            raise NotImplementedError(order)
        return _ceil_cbrt((ndofs / _NP_3D[order]) / 6)
    if dim == 2:
        if order not in _NP_2D:
            raise NotImplementedError(order)
        return int(ceil(sqrt((ndofs / _NP_2D[order]) / 2)))
    raise NotImplementedError


def get_ndof_for_regular_rect_mesh(dim: int, order: int, nel_1d: int) -> int:
    """DOFs per field as the reference counts them (utils.py:46-74).

    Bit-for-bit the reference's integer, including its 3D P1 case, which
    multiplies the 6*n^3 tets by 6 (the tets per hexahedron) where a P1 tet
    has 4 nodes.  The reference's DOF-scaling tool divides by this count
    (tools/wave_dof_scaling.py:158-167), so GDOF/s figures comparable with
    it use it; :func:`count_nodes_regular_rect_mesh` is the true count."""
    if dim == 3:
        if order not in _NP_3D:
            raise NotImplementedError(order)
        return 6 * nel_1d ** 3 * (6 if order == 1 else _NP_3D[order])
    if dim == 2:
        if order not in _NP_2D:
            raise NotImplementedError(order)
        return 2 * nel_1d ** 2 * _NP_2D[order]
    raise NotImplementedError


def count_nodes_regular_rect_mesh(dim: int, order: int, nel_1d: int) -> int:
    """True nodes per field of the box mesh: elements times Np."""
    if dim == 3:
        if order not in _NP_3D:
            raise NotImplementedError(order)
        return 6 * nel_1d ** 3 * _NP_3D[order]
    if dim == 2:
        if order not in _NP_2D:
            raise NotImplementedError(order)
        return 2 * nel_1d ** 2 * _NP_2D[order]
    raise NotImplementedError
