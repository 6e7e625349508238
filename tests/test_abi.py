"""The C-ABI library loads and exports every entry point include/dgfem.h
declares (no compute calls: these run without a GPU)."""

import ctypes
import os
import re

import pytest

from actx_dgfem_suite_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared_functions():
    text = open(os.path.join(ROOT, "include", "dgfem.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(dg_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_the_binding_symbols():
    assert _declared_functions() == sorted(_lib.EXPORTED_SYMBOLS)


@pytest.mark.skipif(not os.path.exists(_lib.LIB_PATH), reason="libdgfem.so not built")
def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(_lib.LIB_PATH)
    for name in _declared_functions():
        assert hasattr(lib, name), name


@pytest.mark.skipif(not os.path.exists(_lib.LIB_PATH), reason="libdgfem.so not built")
def test_last_error_string_starts_empty():
    lib = _lib.load_library()
    assert lib.dg_last_error_string() == b""


@pytest.mark.skipif(not os.path.exists(_lib.LIB_PATH), reason="libdgfem.so not built")
def test_plan_create_rejects_bad_descriptors_without_touching_the_gpu():
    lib = _lib.load_library()
    handle = ctypes.c_void_p()
    assert lib.dg_plan_create(None, ctypes.byref(handle)) == 1
    desc = _lib.DgPlanDesc()
    desc.equation = 7
    desc.order = 1
    desc.nelements = 1
    assert lib.dg_plan_create(ctypes.byref(desc), ctypes.byref(handle)) == 2
    assert b"equation" in lib.dg_last_error_string()
    desc.equation = 0
    desc.order = 5
    assert lib.dg_plan_create(ctypes.byref(desc), ctypes.byref(handle)) == 2
    desc.order = 1
    assert lib.dg_plan_create(ctypes.byref(desc), ctypes.byref(handle)) == 1  # NULL arrays
    assert lib.dg_plan_destroy(None) == 0


def _small_desc(keep):
    """A valid descriptor of a 2^3 x 6 tet P1 mesh with the adjoint data;
    ``keep`` holds the host arrays (private copies) alive."""
    import numpy as np

    from actx_dgfem_suite_amd.discretization import make_box_discretization

    d = make_box_discretization(1, 2)
    arrs = {
        "d_weak": np.array(d.ref.d_weak, dtype=np.float64, order="C"),
        "d_strong": np.array(d.ref.d_strong, dtype=np.float64, order="C"),
        "lift": np.array(d.ref.lift, dtype=np.float64, order="C"),
        "face_nodes": np.array(d.ref.face_nodes, dtype=np.int32, order="C"),
        "geometry": np.array(d.packed_geometry(), dtype=np.float64, order="C"),
        "ext_map": np.array(d.packed_ext_map(), dtype=np.int32, order="C"),
        "mass_inv": np.array(d.ref.mass_inv, dtype=np.float64, order="C"),
        "face_mass": np.array(d.ref.face_mass, dtype=np.float64, order="C"),
    }
    keep.update(arrs)
    desc = _lib.DgPlanDesc()
    desc.equation = 0
    desc.order = 1
    desc.nelements = d.nelements
    for name, a in arrs.items():
        ctype = ctypes.c_int32 if a.dtype == np.int32 else ctypes.c_double
        setattr(desc, name, a.ctypes.data_as(ctypes.POINTER(ctype)))
    desc.params[0] = 1.0
    return desc


@pytest.mark.skipif(not os.path.exists(_lib.LIB_PATH), reason="libdgfem.so not built")
def test_plan_create_checks_adjoint_data_without_touching_the_gpu():
    """Adjoint inputs are validated on the host before any device call."""
    lib = _lib.load_library()
    handle = ctypes.c_void_p()
    keep = {}
    desc = _small_desc(keep)
    desc.face_mass = None   # mass_inv without face_mass
    assert lib.dg_plan_create(ctypes.byref(desc), ctypes.byref(handle)) == 1
    assert b"together" in lib.dg_last_error_string()

    desc = _small_desc(keep)
    fm = keep["face_mass"]
    fnodes = set(keep["face_nodes"][0].tolist())
    off = next(i for i in range(fm.shape[0]) if i not in fnodes)
    fm[off, 0] = 0.5        # couples a node off face 0
    assert lib.dg_plan_create(ctypes.byref(desc), ctypes.byref(handle)) == 1
    assert b"off its face" in lib.dg_last_error_string()

    desc = _small_desc(keep)
    em = keep["ext_map"]
    row, col = next((r, c) for r in range(em.shape[0]) for c in range(em.shape[1])
                    if em[r, c] >= 0 and c % 3 == 0)
    em[row, col] = (em[row, col] + 4 * 7) % (em.shape[0] * 4)   # another element's node
    assert lib.dg_plan_create(ctypes.byref(desc), ctypes.byref(handle)) == 1
    assert b"ext_map" in lib.dg_last_error_string()


def _small_desc_2d(keep):
    """A valid descriptor of a 3^2 x 2 triangle P2 mesh with the adjoint data."""
    import numpy as np

    from actx_dgfem_suite_amd.discretization.triangle import make_square_discretization

    d = make_square_discretization(2, 3)
    arrs = {
        "d_weak": np.array(d.ref.d_weak, dtype=np.float64, order="C"),
        "d_strong": np.array(d.ref.d_strong, dtype=np.float64, order="C"),
        "lift": np.array(d.ref.lift, dtype=np.float64, order="C"),
        "face_nodes": np.array(d.ref.face_nodes, dtype=np.int32, order="C"),
        "geometry": np.array(d.packed_geometry(), dtype=np.float64, order="C"),
        "ext_map": np.array(d.packed_ext_map(), dtype=np.int32, order="C"),
        "mass_inv": np.array(d.ref.mass_inv, dtype=np.float64, order="C"),
        "face_mass": np.array(d.ref.face_mass, dtype=np.float64, order="C"),
    }
    keep.update(arrs)
    desc = _lib.DgPlanDesc()
    desc.equation = 0
    desc.order = 2
    desc.nelements = d.nelements
    desc.dim = 2
    for name, a in arrs.items():
        ctype = ctypes.c_int32 if a.dtype == np.int32 else ctypes.c_double
        setattr(desc, name, a.ctypes.data_as(ctypes.POINTER(ctype)))
    desc.params[0] = 1.0
    return desc


@pytest.mark.skipif(not os.path.exists(_lib.LIB_PATH), reason="libdgfem.so not built")
def test_plan_create_checks_triangle_descriptors_without_touching_the_gpu():
    """dim selects triangles; their plans reject the MFMA kernel, and their
    exterior map and adjoint data are checked with 3 faces of 3 nodes."""
    lib = _lib.load_library()
    handle = ctypes.c_void_p()
    keep = {}
    desc = _small_desc_2d(keep)
    desc.dim = 4
    assert lib.dg_plan_create(ctypes.byref(desc), ctypes.byref(handle)) == 2
    assert b"dim" in lib.dg_last_error_string()
    desc = _small_desc_2d(keep)
    desc.variant = _lib.KERNEL_VARIANTS["mfma"]
    assert lib.dg_plan_create(ctypes.byref(desc), ctypes.byref(handle)) == 2
    assert b"MFMA" in lib.dg_last_error_string()
    desc = _small_desc_2d(keep)
    desc.face_mass = None
    assert lib.dg_plan_create(ctypes.byref(desc), ctypes.byref(handle)) == 1
    assert b"together" in lib.dg_last_error_string()
    desc = _small_desc_2d(keep)
    fm = keep["face_mass"]
    assert fm.shape == (6, 9)      # (Np, 3 faces x Nfp)
    fnodes = set(keep["face_nodes"][2].tolist())
    off = next(i for i in range(6) if i not in fnodes)
    fm[off, 2 * 3] = 0.5           # face 2 couples a node off that edge
    assert lib.dg_plan_create(ctypes.byref(desc), ctypes.byref(handle)) == 1
    assert b"off its face" in lib.dg_last_error_string()
    desc = _small_desc_2d(keep)
    em = keep["ext_map"]
    assert em.shape == (18, 9)
    em[5, 4] = 18 * 6          # one past the last node
    assert lib.dg_plan_create(ctypes.byref(desc), ctypes.byref(handle)) == 1
    assert b"ext_map" in lib.dg_last_error_string()


def test_missing_library_fails_loudly(tmp_path):
    with pytest.raises(RuntimeError, match="no CPU fallback"):
        _lib.load_library.__wrapped__(str(tmp_path / "libdgfem.so"))


def test_engine_refuses_to_run_without_gpu():
    import torch
    if torch.cuda.is_available():
        pytest.skip("GPU present")
    from actx_dgfem_suite_amd.arraycontext import HIPDGFEMArrayContext
    with pytest.raises(RuntimeError, match="GPU"):
        HIPDGFEMArrayContext()


@pytest.mark.skipif(not os.path.exists(_lib.LIB_PATH), reason="libdgfem.so not built")
def test_integration_stub_loads_and_reports_bad_input():
    """The ctypes stub of INTEGRATION.md, executed as written, binds the
    library and raises with the library's message on a bad descriptor
    (dg_plan_create checks it before any device call)."""
    from tests._helpers import discr_pair, stub_namespace
    ns = stub_namespace()
    for name in ("make_plan", "rhs", "rhs_vjp", "destroy_plan"):
        assert callable(ns[name])
    d, _ = discr_pair(1, 1)
    r = d.ref
    with pytest.raises(RuntimeError, match="order 7 not supported"):
        ns["make_plan"]("wave", 7, r.d_weak, r.d_strong, r.lift, r.face_nodes,
                        d.packed_geometry(), d.packed_ext_map(), 1.0)
    ns["destroy_plan"](None)
