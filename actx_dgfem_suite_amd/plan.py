"""Device plan for one DG operator: owns the library plan handle and launches
the RHS and LSERK4 kernels on torch device memory.

torch supplies device memory and the current stream; all arithmetic happens in
``libdgfem.so``.
"""

from __future__ import annotations

import ctypes
from collections.abc import Sequence

import numpy as np
import torch

from actx_dgfem_suite_amd import _lib
from actx_dgfem_suite_amd.discretization import Discretization

NFIELDS = {"wave": 4, "euler": 5, "maxwell": 6}
# triangles: wave (u, v_x, v_y), Euler (mass, energy, momentum[2]),
# Maxwell in TM form (E_z, H_x, H_y)
NFIELDS_2D = {"wave": 3, "euler": 4, "maxwell": 3}


def _dptr(a: np.ndarray, ctype):
    return a.ctypes.data_as(ctypes.POINTER(ctype))


class DGPlan:
    """Plan for ``equation`` on ``discr`` with scalar ``param``.

    ``param`` is the wave speed c (wave), gamma (Euler) or the flux_type alpha
    (Maxwell), matching equations/wave.py:279, equations/euler.py:151 and
    equations/maxwell.py:57-59.  ``variant`` picks the FP64 kernel:
    ``"auto"``/``"mfma"`` (matrix cores) or ``"lds"`` (the baseline).
    Triangle discretizations (``discr.dim == 2``) run the LDS-staged 2D kernel
    (FP64 arithmetic; the mixed precisions store FP32 fields) and its adjoint.
    """

    def __init__(self, discr: Discretization, equation: str, param: float,
                 device: torch.device | int | None = None, variant: str = "auto"):
        if equation not in _lib.EQUATION_IDS:
            raise NotImplementedError(f"equation '{equation}' not supported")
        if variant not in _lib.KERNEL_VARIANTS:
            raise ValueError(f"kernel variant '{variant}' (known: {sorted(_lib.KERNEL_VARIANTS)})")
        self.variant = variant
        if not torch.cuda.is_available():
            raise RuntimeError("DGPlan needs a ROCm GPU (no CPU fallback)")
        self.lib = _lib.load_library()
        self.discr = discr
        self.equation = equation
        self.param = float(param)
        self.dim = int(getattr(discr, "dim", 3))
        if self.dim not in (2, 3):
            raise NotImplementedError(f"dim={self.dim}")
        self.nfields = (NFIELDS if self.dim == 3 else NFIELDS_2D)[equation]
        self.device = torch.device("cuda", torch.cuda.current_device()
                                   if device is None else torch.device(device).index or 0)
        ref = discr.ref
        # element centroids: the plan's locality schedule (a discretization
        # without node coordinates keeps its element order)
        centroids = None
        if getattr(discr, "nodes", None) is not None:
            centroids = discr.nodes.mean(axis=2).T
            if centroids.shape[1] == 2:
                centroids = np.concatenate([centroids, np.zeros((centroids.shape[0], 1))],
                                           axis=1)
        self._host = {
            "d_weak": np.ascontiguousarray(ref.d_weak, dtype=np.float64),
            "d_strong": np.ascontiguousarray(ref.d_strong, dtype=np.float64),
            "lift": np.ascontiguousarray(ref.lift, dtype=np.float64),
            "face_nodes": np.ascontiguousarray(ref.face_nodes, dtype=np.int32),
            "geometry": np.ascontiguousarray(discr.packed_geometry(), dtype=np.float64),
            "ext_map": discr.packed_ext_map(),
            "mass_inv": np.ascontiguousarray(ref.mass_inv, dtype=np.float64),
            "face_mass": np.ascontiguousarray(ref.face_mass, dtype=np.float64),
            "centroids": (None if centroids is None
                          else np.ascontiguousarray(centroids, dtype=np.float64)),
        }
        h = self._host
        desc = _lib.DgPlanDesc()
        desc.equation = _lib.EQUATION_IDS[equation]
        desc.order = discr.order
        desc.nelements = discr.nelements
        desc.d_weak = _dptr(h["d_weak"], ctypes.c_double)
        desc.d_strong = _dptr(h["d_strong"], ctypes.c_double)
        desc.lift = _dptr(h["lift"], ctypes.c_double)
        desc.face_nodes = _dptr(h["face_nodes"], ctypes.c_int32)
        desc.geometry = _dptr(h["geometry"], ctypes.c_double)
        desc.ext_map = _dptr(h["ext_map"], ctypes.c_int32)
        desc.mass_inv = _dptr(h["mass_inv"], ctypes.c_double)
        desc.face_mass = _dptr(h["face_mass"], ctypes.c_double)
        desc.dim = self.dim
        if h["centroids"] is not None:
            desc.centroids = _dptr(h["centroids"], ctypes.c_double)
        desc.params[0] = self.param
        desc.variant = _lib.KERNEL_VARIANTS[variant]
        handle = ctypes.c_void_p()
        with torch.cuda.device(self.device):
            _lib.check(self.lib, self.lib.dg_plan_create(ctypes.byref(desc),
                                                         ctypes.byref(handle)),
                       "dg_plan_create")
        self._handle = handle
        self._vjp_workspace: dict[torch.dtype, list[torch.Tensor]] = {}
        del self._host  # the plan holds device copies

    @property
    def kernel_variant(self) -> str:
        """The FP64 kernel this plan runs: ``"mfma"`` or ``"lds"``."""
        v = ctypes.c_int32()
        _lib.check(self.lib, self.lib.dg_plan_variant(self._handle, ctypes.byref(v)),
                   "dg_plan_variant")
        return {_lib.DG_KERNEL_MFMA: "mfma", _lib.DG_KERNEL_LDS: "lds"}[v.value]

    # {{{ validation

    @property
    def shape(self) -> tuple[int, int]:
        return (self.discr.nelements, self.discr.np)

    def _check_fields(self, fields: Sequence[torch.Tensor], what: str,
                      dtype: torch.dtype = torch.float64) -> None:
        if len(fields) != self.nfields:
            raise ValueError(f"{what}: expected {self.nfields} fields, got {len(fields)}")
        for k, t in enumerate(fields):
            if not isinstance(t, torch.Tensor):
                raise TypeError(f"{what}[{k}]: expected torch.Tensor, got {type(t)}")
            if t.device != self.device:
                raise ValueError(f"{what}[{k}] on {t.device}, plan on {self.device}")
            if t.dtype != dtype:
                raise TypeError(f"{what}[{k}]: expected {dtype}, got {t.dtype}")
            if tuple(t.shape) != self.shape or not t.is_contiguous():
                raise ValueError(f"{what}[{k}]: expected contiguous {self.shape}, "
                                 f"got {tuple(t.shape)}")

    # }}}

    def empty_fields(self, dtype: torch.dtype = torch.float64) -> list[torch.Tensor]:
        return [torch.empty(self.shape, dtype=dtype, device=self.device)
                for _ in range(self.nfields)]

    def rhs(self, q: Sequence[torch.Tensor], out: Sequence[torch.Tensor] | None = None,
            precision: str = "fp64") -> list[torch.Tensor]:
        """Launch one RHS evaluation on the current stream; returns ``out``.

        ``precision="fp64"``: float64 fields (the drop-in form).
        ``"fp32"`` / ``"bf16x3"``: the mixed form on float32 fields.
        """
        if precision in _lib.MIXED_PRECISIONS:
            return self.rhs_mixed(q, out, precision)
        if precision not in _lib.PRECISION_IDS:
            raise NotImplementedError(f"precision '{precision}'")
        self._check_fields(q, "q")
        if out is None:
            out = self.empty_fields()
        else:
            self._check_fields(out, "out")
        pin = _lib.pointer_array([t.data_ptr() for t in q])
        pout = _lib.pointer_array([t.data_ptr() for t in out])
        stream = torch.cuda.current_stream(self.device).cuda_stream
        _lib.check(self.lib, self.lib.dg_rhs(self._handle, _lib.PRECISION_IDS[precision],
                                             pin, pout, ctypes.c_void_p(stream)), "dg_rhs")
        return list(out)

    def rhs_mixed(self, q: Sequence[torch.Tensor], out: Sequence[torch.Tensor] | None = None,
                  precision: str = "fp32") -> list[torch.Tensor]:
        """Mixed-precision RHS on float32 fields (``"fp32"`` or ``"bf16x3"``)."""
        if precision not in _lib.MIXED_PRECISIONS:
            raise NotImplementedError(f"mixed precision '{precision}'")
        self._check_fields(q, "q", torch.float32)
        if out is None:
            out = self.empty_fields(torch.float32)
        else:
            self._check_fields(out, "out", torch.float32)
        pin = _lib.pointer_array([t.data_ptr() for t in q])
        pout = _lib.pointer_array([t.data_ptr() for t in out])
        stream = torch.cuda.current_stream(self.device).cuda_stream
        _lib.check(self.lib, self.lib.dg_rhs_mixed(self._handle, _lib.PRECISION_IDS[precision],
                                                   pin, pout, ctypes.c_void_p(stream)),
                   "dg_rhs_mixed")
        return list(out)

    def lserk_stage_mixed(self, a: float, b: float, dt: float, q: Sequence[torch.Tensor],
                          res: Sequence[torch.Tensor], rhs: Sequence[torch.Tensor],
                          q32: Sequence[torch.Tensor]) -> None:
        """FP64 state/residual update from an FP32 RHS; refreshes the FP32 shadow."""
        self._check_fields(q, "q")
        self._check_fields(res, "res")
        self._check_fields(rhs, "rhs", torch.float32)
        self._check_fields(q32, "q32", torch.float32)
        stream = torch.cuda.current_stream(self.device).cuda_stream
        _lib.check(self.lib, self.lib.dg_lserk_stage_mixed(
            self._handle, float(a), float(b), float(dt),
            _lib.pointer_array([t.data_ptr() for t in q]),
            _lib.pointer_array([t.data_ptr() for t in res]),
            _lib.pointer_array([t.data_ptr() for t in rhs]),
            _lib.pointer_array([t.data_ptr() for t in q32]),
            ctypes.c_void_p(stream)), "dg_lserk_stage_mixed")

    def lserk_stage(self, a: float, b: float, dt: float, q: Sequence[torch.Tensor],
                    res: Sequence[torch.Tensor], rhs: Sequence[torch.Tensor]) -> None:
        for name, f in (("q", q), ("res", res), ("rhs", rhs)):
            self._check_fields(f, name)
        stream = torch.cuda.current_stream(self.device).cuda_stream
        _lib.check(self.lib, self.lib.dg_lserk_stage(
            self._handle, float(a), float(b), float(dt),
            _lib.pointer_array([t.data_ptr() for t in q]),
            _lib.pointer_array([t.data_ptr() for t in res]),
            _lib.pointer_array([t.data_ptr() for t in rhs]),
            ctypes.c_void_p(stream)), "dg_lserk_stage")

    @property
    def fused_stage(self) -> bool:
        """Whether ``rhs_lserk_stage`` is available: tetrahedral plans on the
        MFMA kernel."""
        return self.dim == 3 and self.kernel_variant == "mfma"

    def rhs_lserk_stage(self, a: float, b: float, dt: float, q_in: Sequence[torch.Tensor],
                        q_out: Sequence[torch.Tensor], res: Sequence[torch.Tensor]) -> None:
        """One RHS and one LSERK4 stage in one kernel (``dg_rhs_lserk_stage``):
        ``res = a*res + dt*L(q_in); q_out = q_in + b*res``.  ``q_out`` and
        ``res`` must share no memory with ``q_in`` or each other."""
        for name, f in (("q_in", q_in), ("q_out", q_out), ("res", res)):
            self._check_fields(f, name)
        stream = torch.cuda.current_stream(self.device).cuda_stream
        _lib.check(self.lib, self.lib.dg_rhs_lserk_stage(
            self._handle, float(a), float(b), float(dt),
            _lib.pointer_array([t.data_ptr() for t in q_in]),
            _lib.pointer_array([t.data_ptr() for t in q_out]),
            _lib.pointer_array([t.data_ptr() for t in res]),
            ctypes.c_void_p(stream)), "dg_rhs_lserk_stage")

    def rhs_lserk_stage_adjoint(self, a: float, b: float, dt: float, s_in: Sequence[torch.Tensor],
                                s_out: Sequence[torch.Tensor], q_hat: Sequence[torch.Tensor]) -> None:
        """The reverse LSERK4 stage of a linear operator in the variables of
        ``adjoint_space``, in one kernel (``dg_rhs_lserk_stage_adjoint``):
        ``w = L(s_in); q_hat += dt*w; s_out = a*s_in + b*q_hat`` (float64;
        plans with ``fused_stage``)."""
        for name, f in (("s_in", s_in), ("s_out", s_out), ("q_hat", q_hat)):
            self._check_fields(f, name)
        stream = torch.cuda.current_stream(self.device).cuda_stream
        _lib.check(self.lib, self.lib.dg_rhs_lserk_stage_adjoint(
            self._handle, float(a), float(b), float(dt),
            _lib.pointer_array([t.data_ptr() for t in s_in]),
            _lib.pointer_array([t.data_ptr() for t in s_out]),
            _lib.pointer_array([t.data_ptr() for t in q_hat]),
            ctypes.c_void_p(stream)), "dg_rhs_lserk_stage_adjoint")

    @property
    def fused_stage_mixed(self) -> bool:
        """Whether ``rhs_lserk_stage_mixed`` is available: tetrahedral plans."""
        return self.dim == 3

    def rhs_lserk_stage_mixed(self, precision: str, a: float, b: float, dt: float,
                              q32_in: Sequence[torch.Tensor], q32_out: Sequence[torch.Tensor],
                              q: Sequence[torch.Tensor], res: Sequence[torch.Tensor]) -> None:
        """One mixed RHS and one LSERK4 stage in one kernel
        (``dg_rhs_lserk_stage_mixed``): ``res = a*res + dt*L(q32_in); q +=
        b*res; q32_out = float(q)``, with ``q`` and ``res`` float64 (updated
        in place) and ``q32_in``/``q32_out`` two different float32 shadows."""
        if precision not in _lib.MIXED_PRECISIONS:
            raise NotImplementedError(f"mixed precision '{precision}'")
        for name, f, dt_ in (("q32_in", q32_in, torch.float32), ("q32_out", q32_out, torch.float32),
                             ("q", q, torch.float64), ("res", res, torch.float64)):
            self._check_fields(f, name, dt_)
        stream = torch.cuda.current_stream(self.device).cuda_stream
        _lib.check(self.lib, self.lib.dg_rhs_lserk_stage_mixed(
            self._handle, _lib.PRECISION_IDS[precision], float(a), float(b), float(dt),
            _lib.pointer_array([t.data_ptr() for t in q32_in]),
            _lib.pointer_array([t.data_ptr() for t in q32_out]),
            _lib.pointer_array([t.data_ptr() for t in q]),
            _lib.pointer_array([t.data_ptr() for t in res]),
            ctypes.c_void_p(stream)), "dg_rhs_lserk_stage_mixed")

    def rhs_vjp(self, q: Sequence[torch.Tensor] | None, rhs_bar: Sequence[torch.Tensor],
                out: Sequence[torch.Tensor] | None = None) -> list[torch.Tensor]:
        """Vector-Jacobian product ``q_bar = (dL/dq)^T rhs_bar`` at state ``q``.

        float64 fields use ``dg_rhs_vjp``; float32 fields the FP32 adjoint
        ``dg_rhs_vjp_fp32`` (FP64 arithmetic, FP32 storage).  ``q`` is read by
        Euler only (wave and Maxwell are linear) and may be ``None`` for them.
        ``out`` may be ``rhs_bar`` itself, but not ``q``.  Launches on the
        current stream; the workspace is allocated on the first call per dtype
        and reused."""
        dtype = rhs_bar[0].dtype if len(rhs_bar) else torch.float64
        if dtype not in (torch.float64, torch.float32):
            raise TypeError(f"rhs_vjp: float64 or float32 fields, got {dtype}")
        if q is not None:
            self._check_fields(q, "q", dtype)
        elif self.equation == "euler":
            raise ValueError("rhs_vjp: the Euler adjoint needs the state q")
        self._check_fields(rhs_bar, "rhs_bar", dtype)
        if out is None:
            out = self.empty_fields(dtype)
        else:
            self._check_fields(out, "out", dtype)
        ws = self._vjp_workspace.get(dtype)
        if ws is None:
            ws = self._vjp_workspace[dtype] = self.empty_fields(dtype)
        fn, name = ((self.lib.dg_rhs_vjp, "dg_rhs_vjp") if dtype == torch.float64
                    else (self.lib.dg_rhs_vjp_fp32, "dg_rhs_vjp_fp32"))
        stream = torch.cuda.current_stream(self.device).cuda_stream
        pq = None if q is None else _lib.pointer_array([t.data_ptr() for t in q])
        _lib.check(self.lib, fn(
            self._handle, pq,
            _lib.pointer_array([t.data_ptr() for t in rhs_bar]),
            _lib.pointer_array([t.data_ptr() for t in out]),
            _lib.pointer_array([t.data_ptr() for t in ws]),
            ctypes.c_void_p(stream)), name)
        return list(out)

    def lserk_stage_adjoint(self, a: float, b: float, dt: float, q_bar: Sequence[torch.Tensor],
                            s: Sequence[torch.Tensor],
                            w: Sequence[torch.Tensor] | None) -> None:
        """Reverse sweep of one LSERK4 stage: ``q_bar += dt*w; s = a*s + b*q_bar``
        (float64, or float32 for the FP32 adjoint)."""
        dtype = q_bar[0].dtype if len(q_bar) else torch.float64
        if dtype not in (torch.float64, torch.float32):
            raise TypeError(f"lserk_stage_adjoint: float64 or float32 fields, got {dtype}")
        self._check_fields(q_bar, "q_bar", dtype)
        self._check_fields(s, "s", dtype)
        if w is not None:
            self._check_fields(w, "w", dtype)
        fn, name = ((self.lib.dg_lserk_stage_adjoint, "dg_lserk_stage_adjoint")
                    if dtype == torch.float64
                    else (self.lib.dg_lserk_stage_adjoint_fp32, "dg_lserk_stage_adjoint_fp32"))
        stream = torch.cuda.current_stream(self.device).cuda_stream
        _lib.check(self.lib, fn(
            self._handle, float(a), float(b), float(dt),
            _lib.pointer_array([t.data_ptr() for t in q_bar]),
            _lib.pointer_array([t.data_ptr() for t in s]),
            None if w is None else _lib.pointer_array([t.data_ptr() for t in w]),
            ctypes.c_void_p(stream)), name)

    def adjoint_space(self, fields: Sequence[torch.Tensor], to_adjoint: bool,
                      out: Sequence[torch.Tensor] | None = None) -> list[torch.Tensor]:
        """``R M^{-1} fields`` (``to_adjoint``) or ``R M fields``: the change of
        variables in which the linear operators' transpose is the forward RHS
        (``dg_adjoint_space``, or ``dg_adjoint_space_fp32`` for float32
        fields; wave and Maxwell plans).  ``out`` may be ``fields`` itself."""
        dtype = fields[0].dtype if len(fields) else torch.float64
        if dtype not in (torch.float64, torch.float32):
            raise TypeError(f"adjoint_space: float64 or float32 fields, got {dtype}")
        self._check_fields(fields, "fields", dtype)
        if out is None:
            out = self.empty_fields(dtype)
        else:
            self._check_fields(out, "out", dtype)
        fn, name = ((self.lib.dg_adjoint_space, "dg_adjoint_space") if dtype == torch.float64
                    else (self.lib.dg_adjoint_space_fp32, "dg_adjoint_space_fp32"))
        stream = torch.cuda.current_stream(self.device).cuda_stream
        _lib.check(self.lib, fn(
            self._handle, 1 if to_adjoint else 0,
            _lib.pointer_array([t.data_ptr() for t in fields]),
            _lib.pointer_array([t.data_ptr() for t in out]),
            ctypes.c_void_p(stream)), name)
        return list(out)

    def close(self) -> None:
        h = getattr(self, "_handle", None)
        if h is not None and h.value:
            self.lib.dg_plan_destroy(h)
            self._handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
