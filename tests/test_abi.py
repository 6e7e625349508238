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
