"""ctypes binding of ``libdgfem.so`` (C ABI in ``include/dgfem.h``).

The engine has no CPU fallback: importing a compute entry point without the
built library raises ``RuntimeError`` naming the missing file.
"""

from __future__ import annotations

import ctypes
import os
from functools import lru_cache

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libdgfem.so")

DG_EQ_WAVE, DG_EQ_EULER, DG_EQ_MAXWELL = 0, 1, 2
DG_PREC_FP64, DG_PREC_FP32, DG_PREC_BF16X3 = 0, 1, 2
DG_KERNEL_AUTO, DG_KERNEL_MFMA, DG_KERNEL_LDS = 0, 1, 2
KERNEL_VARIANTS = {"auto": DG_KERNEL_AUTO, "mfma": DG_KERNEL_MFMA, "lds": DG_KERNEL_LDS}

EQUATION_IDS = {"wave": DG_EQ_WAVE, "euler": DG_EQ_EULER, "maxwell": DG_EQ_MAXWELL}
PRECISION_IDS = {"fp64": DG_PREC_FP64, "fp32": DG_PREC_FP32, "bf16x3": DG_PREC_BF16X3}
MIXED_PRECISIONS = ("fp32", "bf16x3")

# every symbol include/dgfem.h declares
EXPORTED_SYMBOLS = (
    "dg_plan_create",
    "dg_plan_destroy",
    "dg_plan_info",
    "dg_plan_variant",
    "dg_rhs",
    "dg_rhs_mixed",
    "dg_lserk_stage",
    "dg_rhs_lserk_stage",
    "dg_rhs_lserk_stage_mixed",
    "dg_adjoint_space_fp32",
    "dg_rhs_lserk_stage_adjoint",
    "dg_lserk_stage_mixed",
    "dg_rhs_vjp",
    "dg_lserk_stage_adjoint",
    "dg_rhs_vjp_fp32",
    "dg_rhs_vjp_noalloc",
    "dg_lserk_stage_adjoint_fp32",
    "dg_adjoint_space",
    "dg_last_error_string",
)


class DgPlanDesc(ctypes.Structure):
    _fields_ = [
        ("equation", ctypes.c_int32),
        ("order", ctypes.c_int32),
        ("nelements", ctypes.c_int64),
        ("d_weak", ctypes.POINTER(ctypes.c_double)),
        ("d_strong", ctypes.POINTER(ctypes.c_double)),
        ("lift", ctypes.POINTER(ctypes.c_double)),
        ("face_nodes", ctypes.POINTER(ctypes.c_int32)),
        ("geometry", ctypes.POINTER(ctypes.c_double)),
        ("ext_map", ctypes.POINTER(ctypes.c_int32)),
        ("params", ctypes.c_double * 4),
        ("variant", ctypes.c_int32),
        ("dim", ctypes.c_int32),
        ("mass_inv", ctypes.POINTER(ctypes.c_double)),
        ("face_mass", ctypes.POINTER(ctypes.c_double)),
        ("centroids", ctypes.POINTER(ctypes.c_double)),
    ]


_PP = ctypes.POINTER(ctypes.c_void_p)


@lru_cache(maxsize=None)
def load_library(path: str = LIB_PATH) -> ctypes.CDLL:
    if not os.path.exists(path):
        raise RuntimeError(
            f"libdgfem.so not found at {path}; build it with "
            "`python -c 'import __graft_entry__ as g; g.build()'` "
            "(the DG engine has no CPU fallback)")
    lib = ctypes.CDLL(path)
    lib.dg_plan_create.argtypes = [ctypes.POINTER(DgPlanDesc), ctypes.POINTER(ctypes.c_void_p)]
    lib.dg_plan_create.restype = ctypes.c_int
    lib.dg_plan_destroy.argtypes = [ctypes.c_void_p]
    lib.dg_plan_destroy.restype = ctypes.c_int
    lib.dg_plan_info.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int32),
                                 ctypes.POINTER(ctypes.c_int32),
                                 ctypes.POINTER(ctypes.c_int32),
                                 ctypes.POINTER(ctypes.c_int64)]
    lib.dg_plan_info.restype = ctypes.c_int
    lib.dg_plan_variant.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int32)]
    lib.dg_plan_variant.restype = ctypes.c_int
    lib.dg_rhs.argtypes = [ctypes.c_void_p, ctypes.c_int32, _PP, _PP, ctypes.c_void_p]
    lib.dg_rhs.restype = ctypes.c_int
    lib.dg_lserk_stage.argtypes = [ctypes.c_void_p, ctypes.c_double, ctypes.c_double,
                                   ctypes.c_double, _PP, _PP, _PP, ctypes.c_void_p]
    lib.dg_lserk_stage.restype = ctypes.c_int
    lib.dg_rhs_mixed.argtypes = [ctypes.c_void_p, ctypes.c_int32, _PP, _PP, ctypes.c_void_p]
    lib.dg_rhs_mixed.restype = ctypes.c_int
    lib.dg_lserk_stage_mixed.argtypes = [ctypes.c_void_p, ctypes.c_double, ctypes.c_double,
                                         ctypes.c_double, _PP, _PP, _PP, _PP, ctypes.c_void_p]
    lib.dg_lserk_stage_mixed.restype = ctypes.c_int
    lib.dg_rhs_vjp.argtypes = [ctypes.c_void_p, _PP, _PP, _PP, _PP, ctypes.c_void_p]
    lib.dg_rhs_vjp.restype = ctypes.c_int
    lib.dg_lserk_stage_adjoint.argtypes = [ctypes.c_void_p, ctypes.c_double, ctypes.c_double,
                                           ctypes.c_double, _PP, _PP, _PP, ctypes.c_void_p]
    lib.dg_lserk_stage_adjoint.restype = ctypes.c_int
    lib.dg_rhs_vjp_fp32.argtypes = [ctypes.c_void_p, _PP, _PP, _PP, _PP, ctypes.c_void_p]
    lib.dg_rhs_vjp_fp32.restype = ctypes.c_int
    lib.dg_rhs_vjp_noalloc.argtypes = [ctypes.c_void_p, _PP, _PP, _PP, ctypes.c_void_p]
    lib.dg_rhs_vjp_noalloc.restype = ctypes.c_int
    lib.dg_lserk_stage_adjoint_fp32.argtypes = [ctypes.c_void_p, ctypes.c_double,
                                                ctypes.c_double, ctypes.c_double, _PP, _PP,
                                                _PP, ctypes.c_void_p]
    lib.dg_lserk_stage_adjoint_fp32.restype = ctypes.c_int
    lib.dg_rhs_lserk_stage.argtypes = [ctypes.c_void_p, ctypes.c_double, ctypes.c_double,
                                       ctypes.c_double, _PP, _PP, _PP, ctypes.c_void_p]
    lib.dg_rhs_lserk_stage.restype = ctypes.c_int
    lib.dg_rhs_lserk_stage_mixed.argtypes = [ctypes.c_void_p, ctypes.c_int32, ctypes.c_double,
                                             ctypes.c_double, ctypes.c_double, _PP, _PP, _PP, _PP,
                                             ctypes.c_void_p]
    lib.dg_rhs_lserk_stage_mixed.restype = ctypes.c_int
    lib.dg_adjoint_space.argtypes = [ctypes.c_void_p, ctypes.c_int32, _PP, _PP, ctypes.c_void_p]
    lib.dg_adjoint_space.restype = ctypes.c_int
    lib.dg_adjoint_space_fp32.argtypes = [ctypes.c_void_p, ctypes.c_int32, _PP, _PP,
                                          ctypes.c_void_p]
    lib.dg_adjoint_space_fp32.restype = ctypes.c_int
    lib.dg_rhs_lserk_stage_adjoint.argtypes = [ctypes.c_void_p, ctypes.c_double, ctypes.c_double,
                                               ctypes.c_double, _PP, _PP, _PP, ctypes.c_void_p]
    lib.dg_rhs_lserk_stage_adjoint.restype = ctypes.c_int
    lib.dg_last_error_string.argtypes = []
    lib.dg_last_error_string.restype = ctypes.c_char_p
    return lib


def check(lib: ctypes.CDLL, rc: int, what: str) -> None:
    if rc != 0:
        msg = lib.dg_last_error_string().decode(errors="replace")
        if rc == 2:
            raise NotImplementedError(f"{what}: {msg}")
        if rc == 1:
            raise ValueError(f"{what}: {msg}")
        raise RuntimeError(f"{what} failed (status {rc}): {msg}")


def pointer_array(ptrs) -> ctypes.Array:
    arr = (ctypes.c_void_p * len(ptrs))()
    for k, p in enumerate(ptrs):
        arr[k] = p
    return arr
